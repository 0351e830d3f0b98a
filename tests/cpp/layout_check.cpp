// CPU check of the working-storage layout arithmetic (blochlib_amd/csrc/blfid_layout.h): for
// plan shapes like the engine's (8-spin and 6-spin clusters at 128 steps / 1024 points, rosym
// and general rho, complex64 and FP64, the static and pulse-program kinds) and for every
// available byte count in a band around each batch-size threshold, the chosen batch fits and is
// the largest that fits.  Prints OK and the number of cases.
#include <cstdio>
#include <cstdlib>

#include "../../blochlib_amd/csrc/blfid_layout.h"

using namespace blfid_layout;

static int fails = 0;
static long cases = 0;

static void check(const Shape &s, long K, double avail) {
    cases++;
    const long kb = choose_batch(s, K, avail);
    if (kb == 0) {
        if ((double)layout_bytes(s, 1) <= avail) {
            std::printf("FAIL: no batch although one crystallite fits (avail %.0f)\n", avail);
            fails++;
        }
        return;
    }
    if (kb < 1 || kb > (K > 1 ? K : 1) || (double)layout_bytes(s, kb) > avail) {
        std::printf("FAIL: batch %ld needs %zu > avail %.0f\n", kb, layout_bytes(s, kb), avail);
        fails++;
    }
    if (kb < K && (double)layout_bytes(s, kb + 1) <= avail) {
        std::printf("FAIL: batch %ld is not the largest (avail %.0f)\n", kb, avail);
        fails++;
    }
    // region offsets are 256-byte aligned and in order
    size_t off[R_COUNT];
    const size_t tot = layout_bytes(s, kb, off);
    for (int i = 0; i < R_COUNT; i++)
        if (off[i] % 256 || (i && off[i] < off[i - 1]) || off[i] > tot) {
            std::printf("FAIL: region %d offset %zu\n", i, off[i]);
            fails++;
        }
}

int main() {
    Shape shapes[6];
    // 8-spin 1H cluster: blocks 70, 56, 56, 28, 28, 16 (two packed 8s), 2; rho = Ix, det = I+
    shapes[0].prec = 1; shapes[0].kind = 0; shapes[0].n = 128; shapes[0].npts = 1024;
    shapes[0].N = 256; shapes[0].bmax = 70; shapes[0].rt_from_q = 1; shapes[0].nparts = 12;
    shapes[0].exp_slots = 7086; shapes[0].qr_slots = 10629; shapes[0].Lsum = 13000;
    shapes[0].Estr = 11440; shapes[0].K = 3722;
    // 6-spin cluster, complex64 and FP64, rosym and general rho
    shapes[1] = shapes[0]; shapes[1].N = 64; shapes[1].bmax = 20; shapes[1].Lsum = 952;
    shapes[1].Estr = 800; shapes[1].nparts = 1; shapes[1].exp_slots = 16384;
    shapes[1].qr_slots = 16384;
    shapes[2] = shapes[1]; shapes[2].rt_from_q = 0; shapes[2].prec = 0;
    // 10 spins with blocks beyond the LDS eigen-solver
    shapes[3] = shapes[0]; shapes[3].N = 1024; shapes[3].bmax = 252; shapes[3].n = 64;
    shapes[3].npts = 512; shapes[3].Lsum = 184756; shapes[3].Estr = 167968; shapes[3].nparts = 83;
    shapes[3].eig_slots = 984; shapes[3].K = 1154;
    // static and pulse-program kinds
    shapes[4] = shapes[1]; shapes[4].kind = 1; shapes[4].n = 1; shapes[4].prec = 0;
    shapes[5] = shapes[1]; shapes[5].kind = 2; shapes[5].n = 256; shapes[5].npts = 2048;
    shapes[5].prec = 0; shapes[5].K = 1154;
    for (const Shape &s : shapes) {
        const long K = s.K;
        const size_t fixed = layout_bytes(s, 0);
        const size_t per = layout_bytes(s, 1) - fixed;
        // below one crystallite, at it, and bands around the thresholds of a sweep of batch sizes
        for (long d = -4096; d <= 4096; d += 97) check(s, K, (double)layout_bytes(s, 1) + d);
        for (long kb = 2; kb <= K + 3; kb += (kb < 40 ? 1 : 37)) {
            const double at = (double)layout_bytes(s, kb);
            for (long d = -2048; d <= 2048; d += 61) check(s, K, at + d);
            check(s, K, at);
            check(s, K, at - 1);
        }
        // the linear estimate of round one: fixed + 4096 + k * per with 1000 crystallites
        for (long k = 1; k < 2000; k += 13) check(s, K, (double)fixed + 4096.0 + (double)k * per);
        check(s, 0, (double)layout_bytes(s, 1));
    }
    if (fails) {
        std::printf("%d failures in %ld cases\n", fails, cases);
        return 1;
    }
    std::printf("OK %ld cases\n", cases);
    return 0;
}
