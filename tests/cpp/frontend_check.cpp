// Exercises the C++ front end (include/blfid/solid.hpp) the way a BlochLib program would.
//   frontend_check cpu            host-side checks; engine calls must fail with a typed error
//   frontend_check gpu <crystals> prints the config-1 FID of a ZCW subset, one "re im" per line
//   frontend_check ptop <crystals> prints a 13C-1H-1H ptop FID under 1H phase-alternating rf
#include <blfid/solid.hpp>

#include <cstdio>
#include <cstring>

using namespace blfid;

static int fails = 0;
#define EXPECT(c)                                                                  \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::fprintf(stderr, "FAILED: %s (line %d)\n", #c, __LINE__);          \
            fails++;                                                               \
        }                                                                          \
    } while (0)

static double maxdiff(const matrix &a, const matrix &b) {
    double m = 0;
    for (int i = 0; i < a.rows(); i++)
        for (int j = 0; j < a.cols(); j++) m = std::max(m, std::abs(a(i, j) - b(i, j)));
    return m;
}

static const char *HET =
    "numspin 3\nT 13C 0\nT 1H 1\nT 1H 2\nC 200 4000 0.3 0 10 20 30\nC 0 3000 0.6 1 40 50 60\n"
    "D -20000 0 1 0 30 40\nD -15000 0 2 0 70 10\nD -25000 1 2 0 50 80\nJ 140 0 1\n";

// 32 steps, 1H at 60 kHz alternating 0 / 180 degrees every 4 steps, 13C at 7 kHz, 17 degrees,
// 300 Hz offset on the second half of the period
static pulseSequence het_sequence() {
    pulseSequence seq(32);
    for (int k = 0; k < 32; k += 4) seq.pulse("1H", k, k + 4, 60000.0, (k / 4) % 2 ? 180.0 : 0.0);
    seq.pulse("13C", 16, 32, 7000.0, 17.0, 300.0);
    return seq;
}

static const char *SPINS =
    "numspin 2\nT 1H 0\nT 1H 1\nC 1000 5000 0.5 0 10 20 30\nC -1000 5000 0.5 1 40 50 60\n"
    "D -20000 0 1\n";

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    SolidSys sys = SolidSys::fromText(SPINS);
    if (!std::strcmp(argv[1], "cpu")) {
        EXPECT(sys.spins() == 2 && sys.HS() == 4);
        // [Ix, Iy] = i Iz and Ip = Ix + i Iy on every spin
        for (int s = 0; s < 2; s++) {
            matrix c = sys.Ix(s) * sys.Iy(s) - sys.Iy(s) * sys.Ix(s);
            EXPECT(maxdiff(c, complex(0, 1) * sys.Iz(s)) < 1e-15);
            EXPECT(maxdiff(sys.Ip(s), sys.Ix(s) + complex(0, 1) * sys.Iy(s)) < 1e-15);
        }
        EXPECT(maxdiff(sys.op("Ix_0+2*Iz_1-Ip"), sys.Ix(0) + 2.0 * sys.Iz(1) - sys.Fp()) < 1e-15);
        // Tr(Fx Fp) / Tr(Fp Fm) = 1/2
        EXPECT(std::abs(trace(sys.Fx(), sys.Fp()) / trace(sys.Fp(), adjoint(sys.Fp())) - 0.5) < 1e-15);
        bool threw = false;
        try {
            SolidSys::fromText("numspin 1\nC 0 100 1.5 0\n");
        } catch (BL_exception &) {
            threw = true;
        }
        EXPECT(threw);
        if (argc >= 3) {
            powder p(std::string(argv[2]) + "/ZCW_3_1154");
            double w = 0;
            for (int k = 0; k < p.size(); k++) w += p.weight(k);
            EXPECT(p.size() == 1154 && std::abs(w - 1.0) < 1e-12);
        }
        // without a device the engine reports BLFID_ERR_DEVICE instead of falling back
        oneFID<SolidSys> f(sys, 16, 640000, 10000, 54.7356);
        int status = -1;
        try {
            f.FID(sys.Fx(), sys.Fp(), 0.1, 0.2, 0.3);
            status = 0;
        } catch (BL_exception &e) {
            status = e.status();
        }
        std::printf("engine status without device: %d\n", status);
        SolidSys het = SolidSys::fromText(HET);
        EXPECT(het.isotopeId("13C") == 0 && het.isotopeId("1H") == 1 && het.isotopeId("15N") == -1);
        pulseSequence seq = het_sequence();
        std::vector<blfid_rf> rf = seq.abi(het);
        EXPECT(rf.size() == 2 && rf[0].isotope == 1 && rf[1].isotope == 0);
        EXPECT(rf[0].amp_hz[5] == 60000.0 && std::abs(rf[0].phase_rad[5] - M_PI) < 1e-15);
        EXPECT(rf[1].amp_hz[15] == 0.0 && rf[1].amp_hz[16] == 7000.0 && rf[1].offset_hz[16] == 300.0);
        bool bad = false;
        try {
            seq.pulse("1H", 30, 33, 1.0, 0.0);
        } catch (BL_exception &) {
            bad = true;
        }
        EXPECT(bad);
        std::printf("%s\n", fails ? "FAIL" : "OK");
        return fails ? 1 : 0;
    }
    if (!std::strcmp(argv[1], "gpu") && argc >= 3) {
        powder all(std::string(argv[2]) + "/ZCW_3_1154"), sub;
        for (int k = 0; k < all.size(); k += 29)
            sub.add(all.theta(k), all.phi(k), all.gamma(k), all.weight(k));
        oneFID<SolidSys> f(sys, 256, 640000, 10000, 54.7356);
        std::vector<complex> fid = f.FID(sub, sys.Fx(), sys.Fp());
        std::printf("sw %.17g\n", f.sw());
        for (auto &z : fid) std::printf("%.17g %.17g\n", z.real(), z.imag());
        return 0;
    }
    if (!std::strcmp(argv[1], "ptop") && argc >= 3) {
        powder all(std::string(argv[2]) + "/ZCW_3_1154"), sub;
        for (int k = 7; k < all.size(); k += 163)
            sub.add(all.theta(k), all.phi(k), all.gamma(k), all.weight(k));
        SolidSys het = SolidSys::fromText(HET);
        ptopFID<SolidSys> f(het, 40, 12500.0, 54.7356, het_sequence());
        std::vector<complex> fid = f.FID(sub, het.Ix(0), het.Ip(0));
        for (auto &z : fid) std::printf("%.17g %.17g\n", z.real(), z.imag());
        return 0;
    }
    return 2;
}
