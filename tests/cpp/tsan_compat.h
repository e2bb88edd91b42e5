// Forced into the thread-sanitizer build only: libstdc++ implements condition_variable timed waits on the
// steady clock with pthread_cond_clockwait, which this toolchain's ThreadSanitizer does not intercept
// (it then believes the waiting thread still holds the mutex and reports false races and double locks).
// Without this macro libstdc++ waits with pthread_cond_timedwait, which it does intercept.
#include <bits/c++config.h>
#undef _GLIBCXX_USE_PTHREAD_COND_CLOCKWAIT
