// Exercises the C++ front end (include/blfid/solid.hpp) the way a BlochLib program would.
//   frontend_check cpu            host-side checks; engine calls must fail with a typed error
//   frontend_check gpu <crystals> prints the config-1 FID of a ZCW subset, one "re im" per line
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
    return 2;
}
