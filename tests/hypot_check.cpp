// Host build of the device hypot (fteik_amd/csrc/hypot_borges.h) against
// the C library's hypot on grid-like and random arguments.  Prints the
// number of differing results.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include "../fteik_amd/csrc/hypot_borges.h"

static uint64_t s = 0x9e3779b97f4a7c15ull;
static double u01() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (double) (s >> 11)*(1.0/9007199254740992.0);
}

int main()
{
    long bad = 0, n = 0;
    const double hs[6] = {10.0, 12.5, 20.0, 25.0, 7.3, 250.0};
    for (int r = 0; r < 2000000; r++) {
        double h = (r % 7 == 6) ? 1.0 + 99.0*u01() : hs[r % 6];
        double a = ((double) (1 + (int) (300*u01())) - (1.0 + 300.0*u01()))*h;
        double b = ((double) (1 + (int) (300*u01())) - (1.0 + 300.0*u01()))*h;
        if (r % 5 == 0) b = ((double) (1 + (int) (300*u01())) - 1.0)*h;
        if (r % 11 == 0) a = 0.0;
        n++;
        if (fteik::hypot_borges(a, b) != hypot(a, b)) bad++;
    }
    for (int r = 0; r < 1000000; r++) {
        double a = (u01() - 0.5)*1e6, b = (u01() - 0.5)*1e3;
        n++;
        if (fteik::hypot_borges(a, b) != hypot(a, b)) bad++;
    }
    printf("%ld %ld\n", bad, n);
    return 0;
}
