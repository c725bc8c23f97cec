// Included first in the ThreadSanitizer build of the serving stress test:
// libstdc++ implements condition_variable::wait_for/wait_until on the steady
// clock with pthread_cond_clockwait, which this compiler's ThreadSanitizer
// does not intercept, so it would see the mutex held across the wait and
// report false races and double locks.  Without the macro the library uses
// pthread_cond_timedwait, which it does intercept.
#include <bits/c++config.h>
#undef _GLIBCXX_USE_PTHREAD_COND_CLOCKWAIT
