/*
 * hcm_oracle.c -- clean-room C restatement of the reference chunk mesher (TEST INFRASTRUCTURE ONLY).
 *
 * Restates hc::client::BlockMeshAlgo<Config>::Generate (client/include/client/BlockMeshAlgo.hpp:494-521)
 * for Config = BlockAlgoConfig<int8_t, {0,0,0,16,16,16}, YZX> (client/src/ChunkMeshTask.cpp:113-115),
 * reading blocks and light from 27 neighbour chunks. The block registry is data (orc_registry), loaded
 * from the fixture that oracle/ref_driver.cpp exports from the reference's own property table.
 *
 * Pinning (parity checks that do not merely retype these formulas):
 *   - tests compare this file against the compiled reference (oracle/_ref) on seeded random chunks of
 *     every workload family, and against committed golden fixtures produced by it;
 *   - closed forms: all-air / all-solid -> no quads; a lone stone block -> 6 quads with the worked
 *     vertex words; a full slab -> exactly one top quad;
 *   - invariants: merged quads re-expand to the brute-force visible-face set, quads stay in the chunk,
 *     the AABB is the union of the quad corners.
 * Only tests/, the smoke check and bench.py's CPU-baseline leg may load this library.
 */
#include "hcm_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------------ */
/* registry loading                                                                                   */
/* ------------------------------------------------------------------------------------------------ */

int orc_registry_load(const char *path, orc_registry *reg) {
	memset(reg, 0, sizeof(*reg));
	FILE *fp = fopen(path, "rb");
	if (!fp)
		return ORC_ERR_IO;
	char magic[8];
	uint32_t hdr[4];
	int ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, "HCREG01", 8) == 0 && fread(hdr, 4, 4, fp) == 4;
	if (ok) {
		reg->n_rows = hdr[0], reg->n_meshes = hdr[1], reg->n_tex = hdr[2];
		ok = reg->n_tex <= 8192 && fread(reg->row_of, 2, 65536, fp) == 65536;
	}
	if (ok) {
		reg->rows = (orc_row *)calloc(reg->n_rows ? reg->n_rows : 1, sizeof(orc_row));
		reg->meshes = (orc_mesh *)calloc(reg->n_meshes ? reg->n_meshes : 1, sizeof(orc_mesh));
		ok = reg->rows && reg->meshes && fread(reg->rows, sizeof(orc_row), reg->n_rows, fp) == reg->n_rows &&
		     fread(reg->meshes, sizeof(orc_mesh), reg->n_meshes, fp) == reg->n_meshes &&
		     fread(reg->tex_flags, 1, reg->n_tex, fp) == reg->n_tex;
	}
	uint8_t liq[2];
	if (ok)
		ok = fread(liq, 1, 2, fp) == 2 && fread(&reg->liquid_flow_tex, 2, 1, fp) == 1;
	fclose(fp);
	if (!ok) {
		orc_registry_free(reg);
		return ORC_ERR_FORMAT;
	}
	reg->liquid_source_id = liq[0], reg->liquid_flow_id = liq[1];
	for (uint32_t i = 0; i < 65536; ++i)
		if (reg->row_of[i] >= reg->n_rows) {
			orc_registry_free(reg);
			return ORC_ERR_FORMAT;
		}
	return ORC_OK;
}

void orc_registry_free(orc_registry *reg) {
	free(reg->rows);
	free(reg->meshes);
	reg->rows = NULL;
	reg->meshes = NULL;
}

/* ------------------------------------------------------------------------------------------------ */
/* textures (texture/include/texture/BlockTexture.hpp:18-80)                                          */
/* ------------------------------------------------------------------------------------------------ */

#define TEX_ID(t) ((uint16_t)((t)&0x1fffu))    /* GetID, :28 */
#define TEX_TRANS(t) ((uint8_t)((t) >> 13u))   /* GetTransformation, :27 */

static inline int tex_flag(const orc_registry *reg, uint16_t t, uint8_t f) { return (reg->tex_flags[TEX_ID(t)] & f) != 0; }

/* BlockTexture::Show, :68-76 */
static int tex_show(const orc_registry *reg, uint16_t self, uint16_t r) {
	if (TEX_ID(self) == 0 || self == r)
		return 0;
	int st = tex_flag(reg, self, ORC_TEX_TRANSPARENT), rt = tex_flag(reg, r, ORC_TEX_TRANSPARENT);
	if (!st && !rt)
		return 0;
	if (tex_flag(reg, self, ORC_TEX_LIQUID) && TEX_ID(r) != 0)
		return 0;
	return !st || rt || tex_flag(reg, r, ORC_TEX_LIQUID);
}

/* ------------------------------------------------------------------------------------------------ */
/* mesher state                                                                                       */
/* ------------------------------------------------------------------------------------------------ */

#define MESH_VCAP 65536u   /* a BlockMesh never holds more than 65,532 vertices (BlockMeshAlgo.hpp:397-403) */
#define MESH_ICAP 98304u   /* 6 indices per 4 vertices */

typedef struct builder {
	uint32_t nv, ni;
	uint32_t aabb[6];
	uint32_t transparent;
	uint32_t *v; /* 2 words per vertex */
	uint16_t *ix;
} builder;

typedef struct light4 {
	uint8_t sun[4], torch[4];
	uint8_t ao; /* AO4: 2 bits per vertex (BlockMeshAlgo.hpp:19-29) */
} light4;

typedef struct mctx {
	const orc_registry *reg;
	const uint16_t *const *blocks;
	const uint8_t *const *light;
	orc_mesh_out *out;
	int err;
	builder opq, trn; /* m_opaque_mesh_info, m_transparent_mesh_info (:484-485) */
	/* m_dynamic_block_textures (:489-491), keyed by position in [-1,16]^3 */
	uint8_t dyn_has[18 * 18 * 18];
	uint16_t dyn_tex[18 * 18 * 18][6];
} mctx;

static inline uint32_t slot_of(int x, int y, int z) { /* Position.hpp:138-145 */
	static const uint32_t kL[3] = {1, 2, 0};
	int cx = x < 0 ? 0 : (x >= ORC_CHUNK ? 2 : 1), cy = y < 0 ? 0 : (y >= ORC_CHUNK ? 2 : 1),
	    cz = z < 0 ? 0 : (z >= ORC_CHUNK ? 2 : 1);
	return kL[cx] * 9u + kL[cy] * 3u + kL[cz];
}
static inline uint32_t local_of(int x, int y, int z) { /* Chunk.hpp:57-59, Position.hpp:98-100 */
	return (uint32_t)((x + ORC_CHUNK) % ORC_CHUNK) +
	       ((uint32_t)((y + ORC_CHUNK) % ORC_CHUNK) * ORC_CHUNK + (uint32_t)((z + ORC_CHUNK) % ORC_CHUNK)) * ORC_CHUNK;
}
static inline uint16_t get_block(const mctx *c, int x, int y, int z) { return c->blocks[slot_of(x, y, z)][local_of(x, y, z)]; }
static inline uint8_t get_light(const mctx *c, int x, int y, int z) { return c->light[slot_of(x, y, z)][local_of(x, y, z)]; }
static inline const orc_row *row_of(const mctx *c, uint16_t b) { return &c->reg->rows[c->reg->row_of[b]]; }

/* Block::GetVerticalLightPass || GetCollision != kSolid (BlockMeshAlgo.hpp:119,124; Block.hpp:69,73) */
static inline int direct_pass(const orc_row *r) { return r->transparency == 2 || r->collision != 1; }
/* Block::GetIndirectLightPass (Block.hpp:70) */
static inline int indirect_pass(const orc_row *r) { return r->transparency != 0; }

static inline int dyn_index(int x, int y, int z) { return (x + 1) + 18 * ((y + 1) + 18 * (z + 1)); }

/* get_tex lambda, BlockMeshAlgo.hpp:322-326 */
static inline uint16_t get_tex(const mctx *c, int x, int y, int z, int face) {
	if (x >= -1 && x <= 16 && y >= -1 && y <= 16 && z >= -1 && z <= 16) {
		int d = dyn_index(x, y, z);
		if (c->dyn_has[d])
			return c->dyn_tex[d][face];
	}
	return row_of(c, get_block(c, x, y, z))->tex[face];
}

/* ------------------------------------------------------------------------------------------------ */
/* light4_init (BlockMeshAlgo.hpp:48-163)                                                             */
/* ------------------------------------------------------------------------------------------------ */

/* front-cell offset per face (kLookup1v, :74) */
static const int8_t kFront[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
/* side / corner / side cell offsets per face and vertex (kLookup3v, :75-108); values must match exactly */
static const int8_t kCorner[6][4][3][3] = {
    {{{1, 0, -1}, {1, -1, -1}, {1, -1, 0}},
     {{1, 0, -1}, {1, 1, -1}, {1, 1, 0}},
     {{1, 0, 1}, {1, 1, 1}, {1, 1, 0}},
     {{1, 0, 1}, {1, -1, 1}, {1, -1, 0}}},
    {{{-1, 0, -1}, {-1, -1, -1}, {-1, -1, 0}},
     {{-1, 0, 1}, {-1, -1, 1}, {-1, -1, 0}},
     {{-1, 0, 1}, {-1, 1, 1}, {-1, 1, 0}},
     {{-1, 0, -1}, {-1, 1, -1}, {-1, 1, 0}}},
    {{{0, 1, -1}, {-1, 1, -1}, {-1, 1, 0}},
     {{0, 1, 1}, {-1, 1, 1}, {-1, 1, 0}},
     {{0, 1, 1}, {1, 1, 1}, {1, 1, 0}},
     {{0, 1, -1}, {1, 1, -1}, {1, 1, 0}}},
    {{{0, -1, -1}, {-1, -1, -1}, {-1, -1, 0}},
     {{0, -1, -1}, {1, -1, -1}, {1, -1, 0}},
     {{0, -1, 1}, {1, -1, 1}, {1, -1, 0}},
     {{0, -1, 1}, {-1, -1, 1}, {-1, -1, 0}}},
    {{{0, -1, 1}, {-1, -1, 1}, {-1, 0, 1}},
     {{0, -1, 1}, {1, -1, 1}, {1, 0, 1}},
     {{0, 1, 1}, {1, 1, 1}, {1, 0, 1}},
     {{0, 1, 1}, {-1, 1, 1}, {-1, 0, 1}}},
    {{{0, -1, -1}, {-1, -1, -1}, {-1, 0, -1}},
     {{0, 1, -1}, {-1, 1, -1}, {-1, 0, -1}},
     {{0, 1, -1}, {1, 1, -1}, {1, 0, -1}},
     {{0, -1, -1}, {1, -1, -1}, {1, 0, -1}}},
};

static void light4_init(const mctx *c, light4 *l4, int face, int x, int y, int z) {
	int fx = x + kFront[face][0], fy = y + kFront[face][1], fz = z + kFront[face][2];
	const orc_row *fr = row_of(c, get_block(c, fx, fy, fz));
	int dp3 = direct_pass(fr);
	uint8_t fl = get_light(c, fx, fy, fz);
	uint8_t ao_byte = 0;
	for (int v = 0; v < 4; ++v) {
		int ip[3], dp[3];
		int cx[3], cy[3], cz[3];
		for (int b = 0; b < 3; ++b) {
			cx[b] = x + kCorner[face][v][b][0];
			cy[b] = y + kCorner[face][v][b][1];
			cz[b] = z + kCorner[face][v][b][2];
			const orc_row *r = row_of(c, get_block(c, cx[b], cy[b], cz[b]));
			ip[b] = indirect_pass(r);
			dp[b] = direct_pass(r);
		}
		int ao;
		if ((!dp[0] && !dp[2]) || (!dp[1] && !dp3))
			ao = 0;
		else {
			ao = 3 - !dp[0] - !dp[1] - !dp[2] - !dp3;
			if (ao < 0)
				ao = 0;
		}
		ao_byte = (uint8_t)((ao_byte & ~(3u << (2 * v))) | ((unsigned)ao << (2 * v)));

		uint32_t cnt = 1, s = fl & 15u, t = fl >> 4;
		if (ip[0] || ip[2]) {
			for (int b = 0; b < 3; ++b)
				if (ip[b]) {
					uint8_t l = get_light(c, cx[b], cy[b], cz[b]);
					++cnt;
					s += l & 15u;
					t += l >> 4;
				}
		}
		l4->sun[v] = (uint8_t)((s << 2) / cnt);
		l4->torch[v] = (uint8_t)((t << 2) / cnt);
	}
	l4->ao = ao_byte;
}

static inline unsigned ao_get(uint8_t ao, int i) { return (ao >> (2 * i)) & 3u; }

static inline int light4_eq(const light4 *a, const light4 *b) {
	return a->ao == b->ao && memcmp(a->sun, b->sun, 4) == 0 && memcmp(a->torch, b->torch, 4) == 0;
}

/* Light4::GetFlip, :38-43 (sunlight only) */
static int light4_flip(const light4 *l) {
	int d02 = (int)(ao_get(l->ao, 0) + 1) * (l->sun[0] + 1) - (int)(ao_get(l->ao, 2) + 1) * (l->sun[2] + 1);
	int d13 = (int)(ao_get(l->ao, 1) + 1) * (l->sun[1] + 1) - (int)(ao_get(l->ao, 3) + 1) * (l->sun[3] + 1);
	if (d02 < 0)
		d02 = -d02;
	if (d13 < 0)
		d13 = -d13;
	return d02 < d13;
}

/* ------------------------------------------------------------------------------------------------ */
/* output                                                                                             */
/* ------------------------------------------------------------------------------------------------ */

static void builder_reset(builder *b, uint32_t transparent) {
	b->nv = b->ni = 0;
	b->aabb[0] = b->aabb[1] = b->aabb[2] = 0xFFFFFFFFu; /* AABB<uint32_t>{} : min = max(), max = min() */
	b->aabb[3] = b->aabb[4] = b->aabb[5] = 0u;
	b->transparent = transparent;
}

static void flush_mesh(mctx *c, const builder *b) {
	orc_mesh_out *o = c->out;
	if (o->n_meshes >= ORC_MAX_MESHES) {
		c->err = ORC_ERR_CAPACITY;
		return;
	}
	uint32_t vb = 0, ib = 0;
	if (o->n_meshes) {
		uint32_t l = o->n_meshes - 1;
		vb = o->vert_begin[l] + o->n_verts[l];
		ib = o->idx_begin[l] + o->n_idx[l];
	}
	if (vb + b->nv > o->vert_cap || ib + b->ni > o->idx_cap) {
		c->err = ORC_ERR_CAPACITY;
		return;
	}
	uint32_t m = o->n_meshes++;
	o->vert_begin[m] = vb;
	o->n_verts[m] = b->nv;
	o->idx_begin[m] = ib;
	o->n_idx[m] = b->ni;
	memcpy(o->aabb[m], b->aabb, sizeof(b->aabb));
	o->transparent[m] = b->transparent;
	memcpy(o->verts + 2 * (size_t)vb, b->v, 8 * (size_t)b->nv);
	memcpy(o->idx + ib, b->ix, 2 * (size_t)b->ni);
}

/* Choose the pass and apply the 16-bit vertex-count split; returns the (possibly stale) index base
 * exactly as the reference computes it (uint16_t cur_vertex read before the split, :273-279, :397-403). */
static builder *begin_quad(mctx *c, uint16_t tex, uint16_t *cur_vertex) {
	builder *b = tex_flag(c->reg, tex, ORC_TEX_TRANS_PASS) ? &c->trn : &c->opq;
	*cur_vertex = (uint16_t)b->nv;
	if ((uint32_t)*cur_vertex + 4u > 65535u) {
		uint32_t trans = b->transparent;
		flush_mesh(c, b);
		builder_reset(b, trans);
	}
	return b;
}

static inline void aabb_point(builder *b, uint32_t x, uint32_t y, uint32_t z) {
	uint32_t p[3] = {x, y, z};
	for (int i = 0; i < 3; ++i) {
		if (p[i] < b->aabb[i])
			b->aabb[i] = p[i];
		if (p[i] > b->aabb[3 + i])
			b->aabb[3 + i] = p[i];
	}
}

/* BlockVertex constructor, client/include/client/BlockVertex.hpp:10-14 */
static inline void push_vertex(builder *b, uint32_t x10, uint32_t y10, uint32_t z10, uint32_t axis, uint32_t face,
                               uint32_t ao, uint32_t sun, uint32_t torch, uint16_t tex) {
	uint32_t *w = b->v + 2 * (size_t)b->nv++;
	w[0] = x10 | (y10 << 10u) | (z10 << 20u) | (axis << 30u);
	w[1] = ((uint32_t)TEX_ID(tex) - 1u) | ((uint32_t)TEX_TRANS(tex) << 10u) | (face << 13u) | (ao << 16u) |
	       (sun << 18u) | (torch << 24u);
}

static inline void push_indices(builder *b, uint16_t base, const uint8_t pat[6]) {
	for (int i = 0; i < 6; ++i)
		b->ix[b->ni++] = (uint16_t)(base + pat[i]);
}

static const uint8_t kIdxFlip[6] = {0, 1, 2, 0, 2, 3};    /* :440-446 */
static const uint8_t kIdxNoFlip[6] = {1, 2, 3, 0, 1, 3};  /* :453-459 */
static const uint8_t kIdxCustom[6] = {0, 1, 2, 0, 2, 3};  /* :295-301 */

/* ------------------------------------------------------------------------------------------------ */
/* liquid flow dynamic mesh (block/block/public/Liquid.hpp:151-302)                                   */
/* ------------------------------------------------------------------------------------------------ */

static const int8_t kFlowFetch[8][3] = {{-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 0, 0},
                                       {1, 0, 0},   {-1, 0, 1}, {0, 0, 1},  {1, 0, 1}}; /* :291-300 */

static void make_face(orc_face *f, uint8_t axis, uint8_t face, uint16_t tex, const uint8_t v[4][3]) {
	memset(f, 0, sizeof(*f));
	f->axis = axis;
	f->light_face = face;
	f->render_face = face;
	f->texture = tex;
	for (int i = 0; i < 4; ++i) {
		f->v[i].x = v[i][0];
		f->v[i].y = v[i][1];
		f->v[i].z = v[i][2];
		f->v[i].ao = 4;
	}
}

/* get_vh, Liquid.hpp:181-186 */
static uint8_t flow_vh(int8_t nb, const int8_t nhs[8], int a, int b, int c) {
	int8_t m1 = nb > nhs[a] ? nb : nhs[a], m2 = nhs[b] > nhs[c] ? nhs[b] : nhs[c];
	uint8_t mx = (uint8_t)(m1 > m2 ? m1 : m2);
	if (mx != 16 && (nb == 0) + (nhs[a] == 0) + (nhs[b] == 0) + (nhs[c] == 0) >= 3)
		return 0;
	return mx;
}

/* flow_mesh_gen<Config, Variant>; nb_height = get_flow_height(Variant). */
static uint32_t flow_mesh(const mctx *c, int x, int y, int z, uint32_t nb_height, orc_face *faces, uint16_t tex6[6]) {
	const orc_registry *reg = c->reg;
	uint16_t nbk[8];
	int8_t nhs[8];
	for (int i = 0; i < 8; ++i) {
		nbk[i] = get_block(c, x + kFlowFetch[i][0], y + kFlowFetch[i][1], z + kFlowFetch[i][2]);
		uint8_t id = (uint8_t)(nbk[i] & 0xffu), meta = (uint8_t)(nbk[i] >> 8);
		int h;
		if (id == reg->liquid_source_id)
			h = 16;
		else if (id == reg->liquid_flow_id) {
			uint32_t fh = (uint32_t)meta * 2u; /* get_flow_height: max(variant * (16 / 8), 1u) -> uint8_t */
			if (fh < 1u)
				fh = 1u;
			h = (uint8_t)fh;
		} else
			h = id != 0 ? -1 : 0; /* WaterTrait::IsObstacle: id != kAir (Water.hpp:13) */
		nhs[i] = (int8_t)h;
	}
	int8_t nb = (int8_t)nb_height;
	uint8_t h0 = flow_vh(nb, nhs, 0, 1, 3), h1 = flow_vh(nb, nhs, 1, 2, 4), h2 = flow_vh(nb, nhs, 3, 5, 6),
	        h3 = flow_vh(nb, nhs, 4, 6, 7);
	uint16_t ft = reg->liquid_flow_tex;
	uint8_t src = reg->liquid_source_id, flw = reg->liquid_flow_id;
	if (h0 == 16 && h1 == 16 && h2 == 16 && h3 == 16) {
		for (int i = 0; i < 6; ++i)
			tex6[i] = ft;
		return 0;
	}
	uint32_t n = 0;
	{
		const uint8_t v[4][3] = {{0, h0, 0}, {0, h2, 16}, {16, h3, 16}, {16, h1, 0}};
		make_face(&faces[n++], 1, 2, ft, v);
	}
	int left_full = h0 == 16 && h2 == 16, left_empty = h0 == 0 && h2 == 0;
	int right_full = h1 == 16 && h3 == 16, right_empty = h1 == 0 && h3 == 0;
	int front_full = h2 == 16 && h3 == 16, front_empty = h2 == 0 && h3 == 0;
	int back_full = h0 == 16 && h1 == 16, back_empty = h0 == 0 && h1 == 0;
#define NB_LIQ(i) ((uint8_t)(nbk[i] & 0xffu) == src || (uint8_t)(nbk[i] & 0xffu) == flw)
	if (!left_full && !left_empty && !NB_LIQ(3)) {
		const uint8_t v[4][3] = {{0, h0, 0}, {0, 0, 0}, {0, 0, 16}, {0, h2, 16}};
		make_face(&faces[n++], 0, 1, ft, v);
	}
	if (!right_full && !right_empty && !NB_LIQ(4)) {
		const uint8_t v[4][3] = {{16, 0, 0}, {16, h1, 0}, {16, h3, 16}, {16, 0, 16}};
		make_face(&faces[n++], 0, 0, ft, v);
	}
	if (!back_full && !back_empty && !NB_LIQ(1)) {
		const uint8_t v[4][3] = {{0, 0, 0}, {0, h0, 0}, {16, h1, 0}, {16, 0, 0}};
		make_face(&faces[n++], 2, 5, ft, v);
	}
	if (!front_full && !front_empty && !NB_LIQ(6)) {
		const uint8_t v[4][3] = {{0, h2, 16}, {0, 0, 16}, {16, 0, 16}, {16, h3, 16}};
		make_face(&faces[n++], 2, 4, ft, v);
	}
#undef NB_LIQ
	tex6[2] = 0;
	tex6[1] = left_full ? ft : 0;
	tex6[0] = right_full ? ft : 0;
	tex6[4] = front_full ? ft : 0;
	tex6[5] = back_full ? ft : 0;
	tex6[3] = ft;
	return n;
}

/* ------------------------------------------------------------------------------------------------ */
/* custom meshes (BlockMeshAlgo.hpp:198-306)                                                          */
/* ------------------------------------------------------------------------------------------------ */

/* light4_interpolate, :165-196 */
static inline uint32_t lerp16(uint32_t a, uint32_t b, uint32_t t) { return a * (16u - t) + b * t; }
static inline uint32_t ceil4096(uint32_t s) { return s / 4096u + (s % 4096u ? 1u : 0u); }
static uint8_t interp(const uint8_t lo[4], const uint8_t hi[4], uint32_t du, uint32_t dv, uint32_t dw) {
	uint32_t l = lerp16(lerp16(lo[0], lo[1], du), lerp16(lo[3], lo[2], du), dv);
	uint32_t h = lerp16(lerp16(hi[0], hi[1], du), lerp16(hi[3], hi[2], du), dv);
	return (uint8_t)ceil4096(lerp16(l, h, dw));
}

static void custom_pass(mctx *c) {
	/* YZX scan over [-1, 16]^3: a1 = y, a2 = z, a3 = x (BlockAlgo.hpp:30-33, BlockMeshAlgo.hpp:202-211) */
	for (int y = -1; y <= 16; ++y)
		for (int z = -1; z <= 16; ++z)
			for (int x = -1; x <= 16; ++x) {
				uint16_t b = get_block(c, x, y, z);
				const orc_row *r = row_of(c, b);
				if (r->mesh == ORC_NO_MESH)
					continue;
				const orc_mesh *mesh = &c->reg->meshes[r->mesh];
				orc_face dyn_faces[ORC_MAX_FACES];
				const orc_face *faces = mesh->faces;
				uint32_t face_count = mesh->face_count;
				if (mesh->kind == ORC_MESH_LIQUID_FLOW) {
					int d = dyn_index(x, y, z);
					face_count = flow_mesh(c, x, y, z, mesh->flow_height, dyn_faces, c->dyn_tex[d]);
					c->dyn_has[d] = 1;
					faces = dyn_faces;
				}
				if (x < 0 || x >= 16 || y < 0 || y >= 16 || z < 0 || z >= 16)
					continue;

				int pos[3] = {x, y, z};
				uint32_t base[3] = {(uint32_t)x << 4, (uint32_t)y << 4, (uint32_t)z << 4};
				int cur_lf = 255;
				int lf_axis = 0, u_axis = 0, v_axis = 0;
				light4 lo, hi;
				memset(&lo, 0, sizeof(lo));
				memset(&hi, 0, sizeof(hi));
				for (uint32_t i = 0; i < face_count; ++i) {
					const orc_face *f = &faces[i];
					if (f->light_face != cur_lf) {
						cur_lf = f->light_face;
						lf_axis = cur_lf >> 1;
						u_axis = (lf_axis + 1) % 3;
						v_axis = (lf_axis + 2) % 3;
						if (cur_lf & 1) {
							int t = u_axis;
							u_axis = v_axis;
							v_axis = t;
						}
						int d = 1 - ((cur_lf & 1) << 1); /* BlockFaceProceed, BlockFace.hpp:13-17 */
						int op[3] = {x, y, z}, nb[3] = {x, y, z};
						op[lf_axis] -= d;
						nb[lf_axis] += d;
						light4_init(c, &lo, cur_lf, op[0], op[1], op[2]);
						if (indirect_pass(row_of(c, get_block(c, nb[0], nb[1], nb[2]))))
							light4_init(c, &hi, cur_lf, pos[0], pos[1], pos[2]);
						else
							hi = lo;
					}
					uint16_t cur;
					builder *bd = begin_quad(c, f->texture, &cur);
					for (int k = 0; k < 4; ++k) {
						const orc_vertex *vt = &f->v[k];
						uint8_t vp[3] = {vt->x, vt->y, vt->z};
						uint32_t px = base[0] + vt->x, py = base[1] + vt->y, pz = base[2] + vt->z;
						aabb_point(bd, px, py, pz);
						uint32_t du = vp[u_axis], dv = vp[v_axis], dw = vp[lf_axis] < 16 ? vp[lf_axis] : 16;
						if (cur_lf & 1)
							dw = 16u - dw;
						uint8_t lo_ao[4], hi_ao[4];
						for (int q = 0; q < 4; ++q) {
							lo_ao[q] = (uint8_t)ao_get(lo.ao, q);
							hi_ao[q] = (uint8_t)ao_get(hi.ao, q);
						}
						uint8_t ao = vt->ao;
						if (ao == 4)
							ao = interp(lo_ao, hi_ao, du, dv, dw);
						uint8_t sun = interp(lo.sun, hi.sun, du, dv, dw);
						uint8_t torch = interp(lo.torch, hi.torch, du, dv, dw);
						push_vertex(bd, px, py, pz, f->axis, f->render_face, ao, sun, torch, f->texture);
					}
					push_indices(bd, cur, kIdxCustom);
				}
			}
}

/* ------------------------------------------------------------------------------------------------ */
/* regular faces (BlockMeshAlgo.hpp:308-482)                                                          */
/* ------------------------------------------------------------------------------------------------ */

/* (x, y, z) of slice coordinate ax and mask coordinates (u, v) for each axis (BlockAlgo.hpp:30-41, 70-81):
 * axis x: U = y, V = z; axis y: U = z, V = x; axis z: U = y, V = x. */
static inline void to_xyz(int axis, int ax, int u, int v, int *x, int *y, int *z) {
	if (axis == 0)
		*x = ax, *y = u, *z = v;
	else if (axis == 1)
		*x = v, *y = ax, *z = u;
	else
		*x = v, *y = u, *z = ax;
}

static void regular_axis(mctx *c, int axis) {
	uint16_t tmask[2][256];
	light4 lmask[2][256];
	const int u_axis = axis == 1 ? 2 : 1, v_axis = axis == 0 ? 2 : 0;
	/* (U + 1) % 3 == V holds for axes x and y (:384) */
	const int uv_cyclic = ((u_axis + 1) % 3) == v_axis;
	for (int xa = 0; xa <= 16; ++xa) {
		memset(tmask, 0, sizeof(tmask)); /* the reference's masks are all-empty at each slice start */
		int face_mask = 0;
		for (int u = 0; u < 16; ++u)
			for (int v = 0; v < 16; ++v) {
				int x, y, z, nx, ny, nz;
				to_xyz(axis, xa, u, v, &x, &y, &z);
				to_xyz(axis, xa - 1, u, v, &nx, &ny, &nz);
				int fa = axis << 1, fb = fa | 1;
				uint16_t ta = get_tex(c, nx, ny, nz, fa), tb = get_tex(c, x, y, z, fb);
				int cnt = u * 16 + v;
				if (xa != 0 && tex_show(c->reg, ta, tb)) {
					tmask[0][cnt] = ta;
					light4_init(c, &lmask[0][cnt], fa, nx, ny, nz);
					face_mask |= 1;
				}
				if (xa != 16 && tex_show(c->reg, tb, ta)) {
					tmask[1][cnt] = tb;
					light4_init(c, &lmask[1][cnt], fb, x, y, z);
					face_mask |= 2;
				}
			}
		for (int inv = 0; inv < 2; ++inv) {
			if (!(face_mask & (1 << inv)))
				continue;
			uint16_t *tm = tmask[inv];
			light4 *lm = lmask[inv];
			for (int u = 0; u < 16; ++u)
				for (int v = 0; v < 16;) {
					int cnt = u * 16 + v;
					uint16_t qt = tm[cnt];
					if (TEX_ID(qt) == 0) {
						++v;
						continue;
					}
					light4 ql = lm[cnt];
					int width = 1, height;
					while (v + width < 16 && tm[cnt + width] == qt && light4_eq(&lm[cnt + width], &ql))
						++width;
					for (height = 1; u + height < 16; ++height) {
						int ok = 1;
						for (int k = 0; k < width && ok; ++k) {
							int id = cnt + k + height * 16;
							ok = tm[id] == qt && light4_eq(&lm[id], &ql);
						}
						if (!ok)
							break;
					}
					int x, y, z;
					to_xyz(axis, xa, u, v, &x, &y, &z);
					int uvec[3], vvec[3];
					if (inv ^ uv_cyclic) {
						to_xyz(axis, 0, height, 0, &uvec[0], &uvec[1], &uvec[2]);
						to_xyz(axis, 0, 0, width, &vvec[0], &vvec[1], &vvec[2]);
					} else {
						to_xyz(axis, 0, 0, width, &uvec[0], &uvec[1], &uvec[2]);
						to_xyz(axis, 0, height, 0, &vvec[0], &vvec[1], &vvec[2]);
					}
					uint16_t cur;
					builder *bd = begin_quad(c, qt, &cur);
					uint32_t p0[3] = {(uint32_t)x, (uint32_t)y, (uint32_t)z};
					uint32_t p[4][3];
					for (int i = 0; i < 3; ++i) {
						p[0][i] = p0[i];
						p[1][i] = p0[i] + (uint32_t)uvec[i];
						p[2][i] = p0[i] + (uint32_t)uvec[i] + (uint32_t)vvec[i];
						p[3][i] = p0[i] + (uint32_t)vvec[i];
					}
					/* AABB merge of {p, p + u + v} (:404-409) */
					for (int i = 0; i < 3; ++i) {
						uint32_t lo = p[0][i] << 4, hi = p[2][i] << 4;
						if (lo < bd->aabb[i])
							bd->aabb[i] = lo;
						if (hi > bd->aabb[3 + i])
							bd->aabb[3 + i] = hi;
					}
					int face = (axis << 1) | inv;
					for (int k = 0; k < 4; ++k)
						push_vertex(bd, p[k][0] << 4, p[k][1] << 4, p[k][2] << 4, (uint32_t)axis, (uint32_t)face,
						            ao_get(ql.ao, k), ql.sun[k], ql.torch[k], qt);
					push_indices(bd, cur, light4_flip(&ql) ? kIdxFlip : kIdxNoFlip);
					for (int a = 0; a < height; ++a)
						memset(tm + cnt + 16 * a, 0, 2 * (size_t)width);
					v += width;
				}
		}
	}
}

/* ------------------------------------------------------------------------------------------------ */
/* Generate (BlockMeshAlgo.hpp:494-521)                                                               */
/* ------------------------------------------------------------------------------------------------ */

typedef struct scratch {
	mctx c;
	uint32_t v_opq[2 * MESH_VCAP], v_trn[2 * MESH_VCAP];
	uint16_t i_opq[MESH_ICAP], i_trn[MESH_ICAP];
} scratch;

static int mesh_with(scratch *s, const orc_registry *reg, const uint16_t *const blocks[27],
                     const uint8_t *const light[27], orc_mesh_out *out) {
	mctx *c = &s->c;
	c->reg = reg;
	c->blocks = blocks;
	c->light = light;
	c->out = out;
	c->err = 0;
	memset(c->dyn_has, 0, sizeof(c->dyn_has));
	c->opq.v = s->v_opq, c->opq.ix = s->i_opq;
	c->trn.v = s->v_trn, c->trn.ix = s->i_trn;
	builder_reset(&c->opq, 0);
	builder_reset(&c->trn, 1);
	out->n_meshes = 0;

	custom_pass(c);
	regular_axis(c, 0);
	regular_axis(c, 1);
	regular_axis(c, 2);
	if (c->opq.nv)
		flush_mesh(c, &c->opq);
	if (c->trn.nv)
		flush_mesh(c, &c->trn);
	return c->err;
}

int orc_mesh_chunk(const orc_registry *reg, const uint16_t *const blocks[27], const uint8_t *const light[27],
                   orc_mesh_out *out) {
	scratch *s = (scratch *)malloc(sizeof(scratch));
	if (!s)
		return ORC_ERR_CAPACITY;
	int r = mesh_with(s, reg, blocks, light, out);
	free(s);
	return r;
}

/* ------------------------------------------------------------------------------------------------ */
/* many jobs on a thread pool (CPU baseline)                                                          */
/* ------------------------------------------------------------------------------------------------ */

typedef struct pool_arg {
	const orc_registry *reg;
	const uint16_t *blocks;
	const uint8_t *light;
	const uint32_t *nbr;
	uint32_t n_jobs;
	uint64_t *quads_out;
	uint32_t next; /* shared job cursor */
	int err;
} pool_arg;

static void *pool_worker(void *p) {
	pool_arg *a = (pool_arg *)p;
	scratch *s = (scratch *)malloc(sizeof(scratch));
	orc_mesh_out out;
	memset(&out, 0, sizeof(out));
	out.vert_cap = 8u * MESH_VCAP;
	out.idx_cap = 8u * MESH_ICAP;
	out.verts = (uint32_t *)malloc(8 * (size_t)out.vert_cap);
	out.idx = (uint16_t *)malloc(2 * (size_t)out.idx_cap);
	if (!s || !out.verts || !out.idx) {
		__atomic_store_n(&a->err, ORC_ERR_CAPACITY, __ATOMIC_RELAXED);
		free(s), free(out.verts), free(out.idx);
		return NULL;
	}
	for (;;) {
		uint32_t j = __atomic_fetch_add(&a->next, 1u, __ATOMIC_RELAXED);
		if (j >= a->n_jobs)
			break;
		const uint16_t *b[27];
		const uint8_t *l[27];
		for (int k = 0; k < 27; ++k) {
			b[k] = a->blocks + (size_t)a->nbr[27 * (size_t)j + k] * ORC_CHUNK_VOXELS;
			l[k] = a->light + (size_t)a->nbr[27 * (size_t)j + k] * ORC_CHUNK_VOXELS;
		}
		int r = mesh_with(s, a->reg, b, l, &out);
		if (r)
			__atomic_store_n(&a->err, r, __ATOMIC_RELAXED);
		uint64_t q = 0;
		for (uint32_t m = 0; m < out.n_meshes; ++m)
			q += out.n_verts[m] / 4;
		if (a->quads_out)
			a->quads_out[j] = q;
	}
	free(s), free(out.verts), free(out.idx);
	return NULL;
}

int orc_mesh_many(const orc_registry *reg, const uint16_t *blocks, const uint8_t *light, const uint32_t *nbr,
                  uint32_t n_jobs, uint32_t n_threads, uint64_t *quads_out) {
	pool_arg a = {reg, blocks, light, nbr, n_jobs, quads_out, 0, 0};
	if (n_threads < 1)
		n_threads = 1;
	if (n_threads > 256)
		n_threads = 256;
	pthread_t th[256];
	uint32_t started = 0;
	for (uint32_t t = 1; t < n_threads; ++t)
		if (pthread_create(&th[started], NULL, pool_worker, &a) == 0)
			++started;
	pool_worker(&a);
	for (uint32_t t = 0; t < started; ++t)
		pthread_join(th[t], NULL);
	return a.err;
}

/* ------------------------------------------------------------------------------------------------ */
/* light rebuild before meshing (client/src/ChunkMeshTask.cpp:48-111)                                */
/* ------------------------------------------------------------------------------------------------ */
/* The mesh task's LightAlgo is BlockLightAlgo<Config{bound [-1, 17)^3}, Border 14>
 * (client/include/client/ChunkMeshTask.inl:53-56): in_bound_bordered is [-15, 31)^3 and
 * IsBorderLightInterfere compares a level with the L1 distance to the bound (BlockLightAlgo.hpp:25-42). */
static inline uint32_t lw_index(int x, int y, int z) {
	return (uint32_t)(x + 15) + ORC_LIGHT_W * ((uint32_t)(z + 15) + ORC_LIGHT_W * (uint32_t)(y + 15));
}
static inline int lw_in_bound(int x, int y, int z) {
	return x >= -15 && x < 31 && y >= -15 && y < 31 && z >= -15 && z < 31;
}
static inline unsigned axis_dist(int v) { return v < -1 ? (unsigned)(-1 - v) : (v >= 17 ? (unsigned)(v - 17) : 0u); }
static inline int light_interferes(int x, int y, int z, unsigned lvl) { /* BlockLightAlgo.hpp:32-42 */
	return lvl > axis_dist(x) + axis_dist(y) + axis_dist(z);
}

/* FIFO of (position, level) entries (std::queue in BlockLightAlgo.hpp:22); a position is pushed at most
 * once per level it is raised to, so 16 entries per cell bound the queue. */
typedef struct lqueue {
	uint32_t *e;
	uint32_t head, tail, mask;
} lqueue;
static inline void lq_push(lqueue *q, int x, int y, int z, unsigned lvl) {
	q->e[q->tail++ & q->mask] = (uint32_t)(x + 15) | (uint32_t)(y + 15) << 6 | (uint32_t)(z + 15) << 12 | lvl << 18;
}

typedef struct lctx {
	const orc_registry *reg;
	const uint8_t *light_level;
	const uint16_t *const *blocks;
	const int8_t *const *heights;
	uint8_t *win;
} lctx;

static inline uint16_t lblock(const lctx *c, int x, int y, int z) { return c->blocks[slot_of(x, y, z)][local_of(x, y, z)]; }
static inline unsigned wsun(const lctx *c, int x, int y, int z) { return c->win[lw_index(x, y, z)] & 15u; }

/* BlockLightAlgo::PropagateLight (BlockLightAlgo.hpp:44-73) on one channel (shift 0: sunlight, 4: torchlight) */
static void propagate(lctx *c, lqueue *q, int shift) {
	while (q->head != q->tail) {
		const uint32_t e = q->e[q->head++ & q->mask];
		unsigned lvl = e >> 18;
		if (lvl <= 1u)
			continue;
		--lvl;
		const int x = (int)(e & 63u) - 15, y = (int)((e >> 6) & 63u) - 15, z = (int)((e >> 12) & 63u) - 15;
		for (int f = 0; f < 6; ++f) { /* BlockFaceProceed (block/include/block/BlockFace.hpp:15-17) */
			int n[3] = {x, y, z};
			n[f >> 1] += 1 - ((f & 1) << 1);
			if (!lw_in_bound(n[0], n[1], n[2]) || !light_interferes(n[0], n[1], n[2], lvl))
				continue;
			uint8_t *w = &c->win[lw_index(n[0], n[1], n[2])];
			const unsigned cur = (*w >> shift) & 15u;
			if (lvl > cur && indirect_pass(&c->reg->rows[c->reg->row_of[lblock(c, n[0], n[1], n[2])]])) {
				*w = (uint8_t)((*w & ~(15u << shift)) | lvl << shift);
				if (lvl > 1)
					lq_push(q, n[0], n[1], n[2], lvl);
			}
		}
	}
}

static int light_rebuild_with(lctx *c, lqueue *sq, lqueue *tq) {
	sq->head = sq->tail = tq->head = tq->tail = 0;
	/* initial torchlight and sunlight, torchlight seeds (ChunkMeshTask.cpp:49-59) */
	for (int y = -15; y < 31; ++y)
		for (int z = -15; z < 31; ++z)
			for (int x = -15; x < 31; ++x) {
				const unsigned torch = c->light_level[lblock(c, x, y, z)] & 15u;
				const int yl = (y + 16) % 16;
				const unsigned sun = yl >= c->heights[slot_of(x, y, z)][(x + 16) % 16 + 16 * ((z + 16) % 16)] ? 15u : 0u;
				c->win[lw_index(x, y, z)] = (uint8_t)(sun | torch << 4);
				if (light_interferes(x, y, z, torch))
					lq_push(tq, x, y, z, torch);
			}
	/* sunlight seeds, skipping cells whose six neighbours are all full sunlight (:61-87) */
	for (int y = -15; y < 31; ++y)
		for (int z = -15; z < 31; ++z)
			for (int x = -15; x < 31; ++x) {
				const unsigned sun = wsun(c, x, y, z);
				if (!light_interferes(x, y, z, sun))
					continue;
				if ((x == -15 || wsun(c, x - 1, y, z) == 15) && (x == 30 || wsun(c, x + 1, y, z) == 15) &&
				    (z == -15 || wsun(c, x, y, z - 1) == 15) && (z == 30 || wsun(c, x, y, z + 1) == 15) &&
				    (y == -15 || wsun(c, x, y - 1, z) == 15) && (y == 30 || wsun(c, x, y + 1, z) == 15))
					continue;
				lq_push(sq, x, y, z, sun);
			}
	propagate(c, sq, 0); /* :90-100 */
	propagate(c, tq, 4); /* :101-111 */
	return ORC_OK;
}

static void halo_of_window(const uint8_t *win, uint8_t *halo) {
	for (int y = -1; y < 17; ++y)
		for (int z = -1; z < 17; ++z)
			for (int x = -1; x < 17; ++x)
				halo[(x + 1) + ORC_HALO * ((z + 1) + ORC_HALO * (y + 1))] = win[lw_index(x, y, z)];
}

#define LQ_CAP (1u << 21) /* > 16 entries per cell of the 46^3 buffer */

int orc_light_rebuild(const orc_registry *reg, const uint8_t *light_level, const uint16_t *const blocks[27],
                      const int8_t *const heights[27], uint8_t *win, uint8_t *halo) {
	lqueue sq = {(uint32_t *)malloc(4u * LQ_CAP), 0, 0, LQ_CAP - 1}, tq = {(uint32_t *)malloc(4u * LQ_CAP), 0, 0, LQ_CAP - 1};
	if (!sq.e || !tq.e) {
		free(sq.e), free(tq.e);
		return ORC_ERR_CAPACITY;
	}
	lctx c = {reg, light_level, blocks, heights, win};
	int r = light_rebuild_with(&c, &sq, &tq);
	if (halo)
		halo_of_window(win, halo);
	free(sq.e), free(tq.e);
	return r;
}

/* The 27 light arrays the mesher reads, from the rebuilt buffer (cells outside it stay 0; the mesher reads
 * only [-1, 17)^3, ChunkMeshTask.cpp:120-121). */
static void light_arrays_of_window(const uint8_t *win, uint8_t (*light)[ORC_CHUNK_VOXELS]) {
	memset(light, 0, 27u * ORC_CHUNK_VOXELS);
	for (int y = -15; y < 31; ++y)
		for (int z = -15; z < 31; ++z)
			for (int x = -15; x < 31; ++x)
				light[slot_of(x, y, z)][local_of(x, y, z)] = win[lw_index(x, y, z)];
}

typedef struct relight_arg {
	const orc_registry *reg;
	const uint8_t *light_level;
	const uint16_t *blocks;
	const int8_t *heights;
	const uint32_t *nbr;
	uint32_t n_jobs;
	int mesh;
	uint8_t *halo_out;
	uint64_t *quads_out;
	uint32_t next;
	int err;
} relight_arg;

static void *relight_worker(void *p) {
	relight_arg *a = (relight_arg *)p;
	lqueue sq = {(uint32_t *)malloc(4u * LQ_CAP), 0, 0, LQ_CAP - 1}, tq = {(uint32_t *)malloc(4u * LQ_CAP), 0, 0, LQ_CAP - 1};
	uint8_t *win = (uint8_t *)malloc(ORC_LIGHT_WVOX);
	uint8_t(*light)[ORC_CHUNK_VOXELS] = (uint8_t(*)[ORC_CHUNK_VOXELS])malloc(27u * ORC_CHUNK_VOXELS);
	scratch *s = a->mesh ? (scratch *)malloc(sizeof(scratch)) : NULL;
	orc_mesh_out out;
	memset(&out, 0, sizeof(out));
	if (a->mesh) {
		out.vert_cap = 8u * MESH_VCAP;
		out.idx_cap = 8u * MESH_ICAP;
		out.verts = (uint32_t *)malloc(8 * (size_t)out.vert_cap);
		out.idx = (uint16_t *)malloc(2 * (size_t)out.idx_cap);
	}
	if (!sq.e || !tq.e || !win || !light || (a->mesh && (!s || !out.verts || !out.idx))) {
		__atomic_store_n(&a->err, ORC_ERR_CAPACITY, __ATOMIC_RELAXED);
	} else {
		for (;;) {
			uint32_t j = __atomic_fetch_add(&a->next, 1u, __ATOMIC_RELAXED);
			if (j >= a->n_jobs)
				break;
			const uint16_t *b[27];
			const int8_t *h[27];
			for (int k = 0; k < 27; ++k) {
				b[k] = a->blocks + (size_t)a->nbr[27 * (size_t)j + k] * ORC_CHUNK_VOXELS;
				h[k] = a->heights + (size_t)a->nbr[27 * (size_t)j + k] * (ORC_CHUNK * ORC_CHUNK);
			}
			lctx c = {a->reg, a->light_level, b, h, win};
			light_rebuild_with(&c, &sq, &tq);
			if (a->halo_out)
				halo_of_window(win, a->halo_out + (size_t)j * ORC_HALO_VOX);
			if (a->mesh) {
				light_arrays_of_window(win, light);
				const uint8_t *l[27];
				for (int k = 0; k < 27; ++k)
					l[k] = light[k];
				int r = mesh_with(s, a->reg, b, l, &out);
				if (r)
					__atomic_store_n(&a->err, r, __ATOMIC_RELAXED);
				uint64_t q = 0;
				for (uint32_t m = 0; m < out.n_meshes; ++m)
					q += out.n_verts[m] / 4;
				if (a->quads_out)
					a->quads_out[j] = q;
			}
		}
	}
	free(sq.e), free(tq.e), free(win), free(light), free(s), free(out.verts), free(out.idx);
	return NULL;
}

int orc_relight_many(const orc_registry *reg, const uint8_t *light_level, const uint16_t *blocks,
                     const int8_t *heights, const uint32_t *nbr, uint32_t n_jobs, uint32_t n_threads, int mesh,
                     uint8_t *halo_out, uint64_t *quads_out) {
	relight_arg a = {reg, light_level, blocks, heights, nbr, n_jobs, mesh, halo_out, quads_out, 0, 0};
	if (n_threads < 1)
		n_threads = 1;
	if (n_threads > 256)
		n_threads = 256;
	pthread_t th[256];
	uint32_t started = 0;
	for (uint32_t t = 1; t < n_threads; ++t)
		if (pthread_create(&th[started], NULL, relight_worker, &a) == 0)
			++started;
	relight_worker(&a);
	for (uint32_t t = 0; t < started; ++t)
		pthread_join(th[t], NULL);
	return a.err;
}
