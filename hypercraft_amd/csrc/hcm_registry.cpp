// hcm_registry.cpp -- the library's built-in block registry, flattened for the device.
//
// Restates the reference's block registration (block/register/blocks:1-29, block/block/*.hpp,
// block/block/public/{Trait,CrossMesh,InnerSurfaceMesh,Liquid,Water}.hpp) and texture registration
// (texture/register/block_textures:1-63, texture/internal/generated/block_texture_transparency.inl:1-3)
// as data, then resolves all 65,536 Block::m_data values the way the reference's generated table does
// (block/generator/Block.cpp.in:7-32): variant and transform are taken from the meta byte with the
// per-id bit widths and clamped to the last valid value.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/hcm.h"

namespace {

// Texture ids in the order of texture/register/block_textures.
enum Tex : uint16_t {
	kNone = 0, kStone, kCobblestone, kDirt,
	kGrassPlain, kGrassPlainTop, kGrassPlainSide, kGrassSavanna, kGrassSavannaTop, kGrassSavannaSide,
	kGrassTropical, kGrassTropicalTop, kGrassTropicalSide, kGrassBoreal, kGrassBorealTop, kGrassBorealSide,
	kSand, kGravel, kGlass, kSnow, kBlueIce, kIce, kSandstone, kWater,
	kAcaciaLeaves, kAcaciaLog, kAcaciaLogTop, kAcaciaPlank,
	kJungleLeaves, kJungleLog, kJungleLogTop, kJunglePlank,
	kOakLeaves, kOakLog, kOakLogTop, kOakPlank,
	kSpruceLeaves, kSpruceLog, kSpruceLogTop, kSprucePlank,
	kBirchLeaves, kBirchLog, kBirchLogTop, kBirchPlank,
	kDeadBush, kVine, kApple, kCactusTop, kCactusSide, kCactusBottom, kRedMushroom, kBrownMushroom, kGlowstone,
	kTexCount
};

// Per-texture flags, identical to the generated arrays in block_texture_transparency.inl (they derive
// from the PNG alpha channel and the transparent-pass list, texture/generator/block_tex_gen.cpp:25-29,57-72).
constexpr uint8_t T = HCM_TEX_TRANSPARENT, P = HCM_TEX_TRANS_PASS, L = HCM_TEX_LIQUID;
constexpr uint8_t kTexFlags[kTexCount] = {
    T,         0, 0, 0,                       // none, stone, cobblestone, dirt
    T | P,     0, 0, T | P, 0, 0,             // grass plain/savanna (+top, side)
    T | P,     0, 0, T | P, 0, 0,             // grass tropical/boreal
    0,         0, T, 0, 0, T | P, 0, T | P | L, // sand, gravel, glass, snow, blue ice, ice, sandstone, water
    T | P,     0, 0, 0,                       // acacia leaves, log, log top, plank
    T | P,     0, 0, 0,                       // jungle
    T | P,     0, 0, 0,                       // oak
    T | P,     0, 0, 0,                       // spruce
    T | P,     0, 0, 0,                       // birch
    T,         T, T, T, T, T, T, T,           // dead bush, vine, apple, cactus top/side/bottom, mushrooms
    0,                                        // glowstone
};

constexpr uint16_t tex(uint16_t id, uint16_t trans = 0) { return uint16_t(id | (trans << 13)); }
constexpr uint16_t tex_id(uint16_t t) { return t & 0x1fffu; }
constexpr uint16_t tex_xor(uint16_t t, uint16_t bits) { return uint16_t(t ^ (bits << 13)); }
// BlockTexture transformations (texture/include/texture/BlockTexture.hpp:46-66)
constexpr uint16_t kSwapUV = 1, kNegU = 2, kNegV = 4;
constexpr uint16_t rot_cw(uint16_t t) { return tex_xor(t, kSwapUV | kNegU); }
constexpr uint16_t rot_ccw(uint16_t t) { return tex_xor(t, kSwapUV | kNegV); }
constexpr uint16_t swap_uv(uint16_t t) { return tex_xor(t, kSwapUV); }

enum Face : uint8_t { kRight = 0, kLeft, kTop, kBottom, kFront, kBack };
enum Trans : uint8_t { kOpaque = 0, kSemi = 1, kTransp = 2 };
enum Coll : uint8_t { kCollNone = 0, kSolid = 1, kCollWater = 2 };

using Tex6 = std::array<uint16_t, 6>;
constexpr Tex6 same(uint16_t t) { return {t, t, t, t, t, t}; }
// BLOCK_TEXTURE_BOT_SIDE_TOP(b, s, t) = {s, s, t, b, s, s} (block/block/public/Trait.hpp:17-18)
constexpr Tex6 bot_side_top(uint16_t b, uint16_t s, uint16_t t) { return {s, s, t, b, s, s}; }

// BlockProperty::RotateCW (block/include/block/BlockProperty.hpp:21-64), textures only.
Tex6 rotate_property(const Tex6 &t, int axis) {
	if (axis == 0)
		return {rot_cw(t[0]), rot_ccw(t[1]), t[kFront], t[kBack], t[kBottom], t[kTop]};
	if (axis == 1)
		return {t[kBack], t[kFront], rot_cw(t[2]), rot_ccw(t[3]), t[kRight], t[kLeft]};
	return {t[kTop], t[kBottom], swap_uv(t[kLeft]), swap_uv(t[kRight]), rot_cw(t[4]), rot_ccw(t[5])};
}

bool trans_pass(uint16_t t) { return (kTexFlags[tex_id(t)] & HCM_TEX_TRANS_PASS) != 0; }

hcm_mesh_face mk_face(uint8_t axis, uint8_t light_face, uint8_t render_face, uint16_t texture,
                      const std::array<std::array<uint8_t, 3>, 4> &v) {
	hcm_mesh_face f;
	std::memset(&f, 0, sizeof(f));
	f.axis = axis, f.light_face = light_face, f.render_face = render_face, f.texture = texture;
	for (int i = 0; i < 4; ++i)
		f.v[i] = {v[i][0], v[i][1], v[i][2], 4};
	return f;
}

// Face slots past face_count are kept zeroed so equal meshes compare equal byte for byte.
void clear_unused(hcm_custom_mesh &m) {
	for (uint32_t f = m.face_count; f < HCM_MAX_FACES; ++f)
		std::memset(&m.faces[f], 0, sizeof(m.faces[f]));
}

hcm_custom_mesh empty_mesh() {
	hcm_custom_mesh m;
	std::memset(&m, 0, sizeof(m));
	return m;
}

// Two crossed planes (block/block/public/CrossMesh.hpp:7-52); a transparent-pass texture keeps only
// the first two faces.
hcm_custom_mesh cross_mesh(uint16_t tex_id_, int r, int lo, int hi, uint8_t light_face) {
	hcm_custom_mesh m = empty_mesh();
	auto a = uint8_t(8 - r), b = uint8_t(8 + r), l = uint8_t(lo), h = uint8_t(hi);
	uint16_t t0 = tex(tex_id_), t1 = tex(tex_id_, kNegU);
	m.faces[0] = mk_face(0, light_face, kLeft, t0, {{{a, l, a}, {b, l, b}, {b, h, b}, {a, h, a}}});
	m.faces[1] = mk_face(2, light_face, kBack, t1, {{{a, l, b}, {b, l, a}, {b, h, a}, {a, h, b}}});
	m.faces[2] = mk_face(0, light_face, kRight, t0, {{{a, h, a}, {b, h, b}, {b, l, b}, {a, l, a}}});
	m.faces[3] = mk_face(2, light_face, kFront, t1, {{{a, h, b}, {b, h, a}, {b, l, a}, {a, l, b}}});
	m.face_count = trans_pass(t0) ? 2u : 4u;
	clear_unused(m);
	return m;
}

// A double-sided plane parallel to one block face at distance dist
// (block/block/public/InnerSurfaceMesh.hpp:7-43).
hcm_custom_mesh inner_surface_mesh(uint16_t tex_id_, uint8_t face, uint8_t dist) {
	hcm_custom_mesh m = empty_mesh();
	uint8_t axis = face >> 1, u = (axis + 1) % 3, v = (axis + 2) % 3;
	uint8_t du[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, x[3] = {0, 0, 0};
	if (face & 1u) {
		du[v] = dv[u] = 16;
		x[axis] = uint8_t(16 - dist);
	} else {
		dv[v] = du[u] = 16;
		x[axis] = dist;
	}
	auto p = [&](int su, int sv) {
		return std::array<uint8_t, 3>{uint8_t(x[0] + su * du[0] + sv * dv[0]), uint8_t(x[1] + su * du[1] + sv * dv[1]),
		                              uint8_t(x[2] + su * du[2] + sv * dv[2])};
	};
	m.faces[0] = mk_face(axis, face, face, tex(tex_id_), {{p(0, 0), p(1, 0), p(1, 1), p(0, 1)}});
	m.faces[1] = mk_face(axis, face, uint8_t(face ^ 1u), tex(tex_id_), {{p(0, 0), p(0, 1), p(1, 1), p(1, 0)}});
	m.face_count = trans_pass(tex(tex_id_)) ? 1u : 2u;
	clear_unused(m);
	return m;
}

// Four inset side faces (block/block/cactus.hpp:4-48).
hcm_custom_mesh cactus_sides() {
	hcm_custom_mesh m = empty_mesh();
	uint16_t t = tex(kCactusSide);
	m.faces[0] = mk_face(0, kRight, kRight, t, {{{15, 0, 16}, {15, 0, 0}, {15, 16, 0}, {15, 16, 16}}});
	m.faces[1] = mk_face(0, kLeft, kLeft, t, {{{1, 0, 0}, {1, 0, 16}, {1, 16, 16}, {1, 16, 0}}});
	m.faces[2] = mk_face(2, kFront, kFront, t, {{{0, 0, 15}, {16, 0, 15}, {16, 16, 15}, {0, 16, 15}}});
	m.faces[3] = mk_face(2, kBack, kBack, t, {{{16, 0, 1}, {0, 0, 1}, {0, 16, 1}, {16, 16, 1}}});
	m.face_count = 4;
	return m;
}

// The liquid flow mesh is computed per cell from its 8 horizontal neighbours
// (block/block/public/Liquid.hpp:156-302); the registry records the level's own height
// get_flow_height(Variant) = max(Variant * 16 / kFlowDistance, 1) (Liquid.hpp:151-154).
hcm_custom_mesh liquid_flow_mesh(uint32_t variant) {
	hcm_custom_mesh m = empty_mesh();
	m.kind = HCM_MESH_LIQUID_FLOW;
	m.flow_height = std::max<uint32_t>(variant * (16u / 8u), 1u);
	return m;
}

struct Prop {
	Tex6 tex{};
	uint8_t transparency = kOpaque;
	uint8_t collision = kSolid;
	bool has_mesh = false;
	hcm_custom_mesh mesh{};
	uint8_t light_level = 0; // BlockProperty::light_level (block/include/block/BlockProperty.hpp:22)
};

Prop cube(const Tex6 &t, uint8_t transparency = kOpaque, uint8_t collision = kSolid) {
	Prop p;
	p.tex = t, p.transparency = transparency, p.collision = collision;
	return p;
}
Prop meshed(const Tex6 &t, uint8_t transparency, uint8_t collision, const hcm_custom_mesh &m) {
	Prop p = cube(t, transparency, collision);
	p.has_mesh = true, p.mesh = m;
	return p;
}
Prop air_like() { return cube(same(kNone), kTransp, kCollNone); }

// Block id description: variant count, transform count and property(variant, transform)
// (Trait.hpp:24-65 for the bit widths: ceil(log2(count)) bits each, 0 for an absent field).
struct IdDesc {
	uint8_t variants = 0, transforms = 0;
	Prop (*get)(uint8_t variant, uint8_t transform) = nullptr;
};

uint8_t bit_width_of(uint8_t n) { // countr_zero(bit_ceil(n)), with bit_ceil(0) = 1
	uint8_t b = 0;
	while ((1u << b) < n)
		++b;
	return b;
}

const uint16_t kLeaves[5] = {kOakLeaves, kAcaciaLeaves, kJungleLeaves, kSpruceLeaves, kBirchLeaves};
const uint16_t kLogSide[5] = {kOakLog, kAcaciaLog, kJungleLog, kSpruceLog, kBirchLog};
const uint16_t kLogTop[5] = {kOakLogTop, kAcaciaLogTop, kJungleLogTop, kSpruceLogTop, kBirchLogTop};
const uint16_t kPlanks[5] = {kOakPlank, kAcaciaPlank, kJunglePlank, kSprucePlank, kBirchPlank};
const uint16_t kGrassSide[4] = {kGrassPlainSide, kGrassSavannaSide, kGrassTropicalSide, kGrassBorealSide};
const uint16_t kGrassTop[4] = {kGrassPlainTop, kGrassSavannaTop, kGrassTropicalTop, kGrassBorealTop};
const uint16_t kGrassCross[4] = {kGrassPlain, kGrassSavanna, kGrassTropical, kGrassBoreal};

std::array<IdDesc, 256> make_ids() {
	std::array<IdDesc, 256> d{};
	auto single = [&](int id, Prop (*g)(uint8_t, uint8_t)) { d[id] = {1, 0, g}; };
	single(0, [](uint8_t, uint8_t) { return air_like(); });
	single(1, [](uint8_t, uint8_t) { return cube(same(kStone)); });
	single(2, [](uint8_t, uint8_t) { return cube(same(kCobblestone)); });
	single(3, [](uint8_t, uint8_t) { return cube(same(kDirt)); });
	d[4] = {4, 0, [](uint8_t v, uint8_t) { return cube(bot_side_top(kDirt, kGrassSide[v], kGrassTop[v])); }};
	d[5] = {4, 0, [](uint8_t v, uint8_t) {
		        return meshed(same(kNone), kTransp, kCollNone, cross_mesh(kGrassCross[v], 8, 0, 16, kTop));
	        }};
	single(6, [](uint8_t, uint8_t) { return cube(same(kSand)); });
	single(7, [](uint8_t, uint8_t) {
		return meshed(same(kNone), kTransp, kCollNone, cross_mesh(kDeadBush, 8, 0, 16, kTop));
	});
	single(8, [](uint8_t, uint8_t) { return cube(same(kGravel)); });
	single(9, [](uint8_t, uint8_t) { return cube(same(kGlass), kTransp, kSolid); });
	single(10, [](uint8_t, uint8_t) { return cube(same(kSnow)); });
	single(11, [](uint8_t, uint8_t) { return cube(same(kBlueIce)); });
	single(12, [](uint8_t, uint8_t) { return cube(same(kIce), kSemi, kSolid); });
	single(13, [](uint8_t, uint8_t) { return cube(same(kSandstone)); });
	single(14, [](uint8_t, uint8_t) { return cube(same(kWater), kSemi, kCollWater); });
	d[15] = {9, 0, [](uint8_t v, uint8_t) { return meshed(same(kNone), kSemi, kCollWater, liquid_flow_mesh(v)); }};
	d[16] = {5, 0, [](uint8_t v, uint8_t) { return cube(same(kLeaves[v]), kSemi, kSolid); }};
	// Log::TransformProperty: transform 0 rotates about z, 2 about x, 1 keeps the upright log (log.hpp:26-34).
	d[17] = {5, 3, [](uint8_t v, uint8_t t) {
		         Tex6 base = bot_side_top(kLogTop[v], kLogSide[v], kLogTop[v]);
		         return cube(t == 0 ? rotate_property(base, 2) : t == 2 ? rotate_property(base, 0) : base);
	         }};
	d[18] = {5, 0, [](uint8_t v, uint8_t) { return cube(same(kPlanks[v])); }};
	single(19, [](uint8_t, uint8_t) {
		return meshed(same(kNone), kTransp, kCollNone, cross_mesh(kApple, 5, 1, 15, kBottom));
	});
	single(20, [](uint8_t, uint8_t) {
		return meshed(bot_side_top(kCactusBottom, kNone, kCactusTop), kSemi, kSolid, cactus_sides());
	});
	d[21] = {1, 6, [](uint8_t, uint8_t t) {
		         return meshed(same(kNone), kTransp, kCollNone, inner_surface_mesh(kVine, t, 1));
	         }};
	// The mushrooms pass `true` as the light-face argument, i.e. face 1 (red_mushroom.hpp:8, brown_mushroom.hpp:8).
	single(22, [](uint8_t, uint8_t) {
		return meshed(same(kNone), kTransp, kCollNone, cross_mesh(kRedMushroom, 5, 0, 12, 1));
	});
	single(23, [](uint8_t, uint8_t) {
		return meshed(same(kNone), kTransp, kCollNone, cross_mesh(kBrownMushroom, 6, 0, 9, 1));
	});
	single(24, [](uint8_t, uint8_t) { // glowstone.hpp: opaque, light level 15
		Prop p = cube(same(kGlowstone));
		p.light_level = 15;
		return p;
	});
	return d;
}

struct Built {
	std::vector<uint16_t> row_of = std::vector<uint16_t>(65536);
	std::vector<hcm_block_row> rows;
	std::vector<hcm_custom_mesh> meshes;
	std::vector<uint8_t> tex_flags;
	std::vector<uint8_t> light_level = std::vector<uint8_t>(65536);
	hcm_registry reg{};

	Built() {
		auto ids = make_ids();
		std::map<std::vector<uint8_t>, uint16_t> row_index, mesh_index;
		auto bytes = [](const void *p, size_t n) {
			auto c = static_cast<const uint8_t *>(p);
			return std::vector<uint8_t>(c, c + n);
		};
		for (uint32_t data = 0; data < 65536; ++data) {
			const IdDesc &id = ids[data & 0xffu];
			uint8_t meta = uint8_t(data >> 8);
			uint8_t vb = id.variants ? bit_width_of(id.variants) : 0, tb = bit_width_of(id.transforms);
			// kBlockDataVariant / kBlockDataTransform: min(field, count - 1) as uint8 (count 0 -> 255)
			uint8_t v = std::min<uint8_t>(uint8_t(meta & ((1u << vb) - 1u)), uint8_t(id.variants - 1u));
			uint8_t t = std::min<uint8_t>(uint8_t((meta >> vb) & ((1u << tb) - 1u)), uint8_t(id.transforms - 1u));
			Prop p = id.get ? id.get(v, t) : air_like();
			light_level[data] = p.light_level;
			hcm_block_row r;
			std::memset(&r, 0, sizeof(r));
			for (int f = 0; f < 6; ++f)
				r.tex[f] = p.tex[f];
			r.transparency = p.transparency;
			r.collision = p.collision;
			r.mesh = HCM_NO_MESH;
			if (p.has_mesh) {
				auto key = bytes(&p.mesh, sizeof(p.mesh));
				auto it = mesh_index.find(key);
				if (it == mesh_index.end()) {
					it = mesh_index.emplace(key, uint16_t(meshes.size())).first;
					meshes.push_back(p.mesh);
				}
				r.mesh = it->second;
			}
			auto key = bytes(&r, sizeof(r));
			auto it = row_index.find(key);
			if (it == row_index.end()) {
				it = row_index.emplace(key, uint16_t(rows.size())).first;
				rows.push_back(r);
			}
			row_of[data] = it->second;
		}
		tex_flags.assign(kTexFlags, kTexFlags + kTexCount);
		reg.n_rows = uint32_t(rows.size());
		reg.n_meshes = uint32_t(meshes.size());
		reg.n_tex = uint32_t(tex_flags.size());
		reg.row_of = row_of.data();
		reg.rows = rows.data();
		reg.meshes = meshes.data();
		reg.tex_flags = tex_flags.data();
		reg.liquid_source_id = 14; // WaterTrait::kSourceID = kWater (Water.hpp:6)
		reg.liquid_flow_id = 15;   // WaterTrait::kFlowID = kFlowingWater
		reg.liquid_flow_tex = kWater;
		reg.light_level = light_level.data();
	}
};

static_assert(sizeof(hcm_block_row) == 16, "row layout");
static_assert(sizeof(hcm_mesh_face) == 24, "face layout");
static_assert(sizeof(hcm_custom_mesh) == 400, "mesh layout");

} // namespace

extern "C" const hcm_registry *hcm_registry_default(void) {
	static const Built built;
	return &built.reg;
}
