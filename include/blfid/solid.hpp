// blfid/solid.hpp -- C++ front end of the MI355X powder FID engine with BlochLib's API shape.
//
// Header-only; links against libblfid.so (include/blfid.h).  It mirrors the classes a
// BlochLib program uses on the solid-2.0 FID path, so such a program switches by changing its
// includes and namespace:
//
//   BlochLib                                           here
//   SolidSys::read(Vector<string>)  sys.cc:120-204     blfid::SolidSys::read(lines)
//   SpinSys::Ix/Iy/Iz/Ip/Im/Fx/Fp.. multispinop.cc     blfid::SolidSys::Ix(i) ... Fp()
//   HamiltonianGen("Ip_0+Ix") ham_gen.cc:269-360       blfid::SolidSys::op("Ip_0+Ix")
//   powder(file) powder.cc:313-389                     blfid::powder(file)
//   oneFID<SolidSys>(sys, npts, sw, wr, rotb, t0)      blfid::oneFID<SolidSys>(...)
//     oneFID.h:103-108, setters :130-177
//   oneFID::FID(powder&, ro, det)  oneFID.h:200-208    FID(pow, ro, det)  (whole powder on GPU)
//   oneFID::FID(ro, det, th, ph, ga) oneFID.h:210-218  FID(ro, det, th, ph, ga)
//   oneFID::sw() (adjusted n*wr) oneFID.h:159, :429    sw()
//   BL_exception blassert.h:52-79                      blfid::BL_exception
//   solid-2.0 pulse sub-sequence + ptop() + fid()       blfid::pulseSequence + ptopFID
//     PulseData::Pulse pulsedata.cc:110-125,             (one rotor period of equal steps,
//     Propogation::propogator propogation.cc:240-304,    whole powder on GPU)
//     ptop doFID sequencerun_funcs.cc:624-655
//
// Operators are dense complex matrices, column-major, spin 0 the leftmost Kronecker factor.
#pragma once

#include <blfid.h>

#include <cmath>
#include <complex>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace blfid {

typedef std::complex<double> complex;

class BL_exception : public std::runtime_error {
  public:
    explicit BL_exception(const std::string &m, int status = 1)
        : std::runtime_error(m), status_(status) {}
    int status() const { return status_; }

  private:
    int status_;
};

inline void check(blfid_status s) {
    if (s != BLFID_OK) throw BL_exception(blfid_last_error(), (int)s);
}

// Dense complex matrix, column-major (BlochLib FullMatrix storage, fmatrix.h:97-121).
class matrix {
  public:
    matrix() : n_(0) {}
    explicit matrix(int n, complex v = 0.0) : n_(n), a_((size_t)n * n, v) {}
    static matrix identity(int n) {
        matrix m(n);
        for (int i = 0; i < n; i++) m(i, i) = 1.0;
        return m;
    }
    int rows() const { return n_; }
    int cols() const { return n_; }
    complex &operator()(int i, int j) { return a_[i + (size_t)j * n_]; }
    const complex &operator()(int i, int j) const { return a_[i + (size_t)j * n_]; }
    const double *data() const { return reinterpret_cast<const double *>(a_.data()); }
    matrix &operator+=(const matrix &o) {
        for (size_t q = 0; q < a_.size(); q++) a_[q] += o.a_[q];
        return *this;
    }
    matrix &operator-=(const matrix &o) {
        for (size_t q = 0; q < a_.size(); q++) a_[q] -= o.a_[q];
        return *this;
    }
    matrix &operator*=(complex s) {
        for (auto &x : a_) x *= s;
        return *this;
    }
    friend matrix operator+(matrix a, const matrix &b) { return a += b; }
    friend matrix operator-(matrix a, const matrix &b) { return a -= b; }
    friend matrix operator*(complex s, matrix a) { return a *= s; }
    friend matrix operator*(const matrix &a, const matrix &b) {
        matrix c(a.n_);
        for (int j = 0; j < a.n_; j++)
            for (int k = 0; k < a.n_; k++) {
                complex bk = b(k, j);
                if (bk == 0.0) continue;
                for (int i = 0; i < a.n_; i++) c(i, j) += a(i, k) * bk;
            }
        return c;
    }
    friend matrix adjoint(const matrix &a) {
        matrix c(a.n_);
        for (int i = 0; i < a.n_; i++)
            for (int j = 0; j < a.n_; j++) c(j, i) = std::conj(a(i, j));
        return c;
    }
    // Tr(A B) = sum_ij A_ij B_ji (matmat.h:1222-1238)
    friend complex trace(const matrix &a, const matrix &b) {
        complex t = 0;
        for (int i = 0; i < a.n_; i++)
            for (int j = 0; j < a.n_; j++) t += a(i, j) * b(j, i);
        return t;
    }

  private:
    int n_;
    std::vector<complex> a_;
};

// Tensor-described SolidSys: spin-1/2 isotopes plus C / D / J tables.
class SolidSys {
  public:
    SolidSys() : names_(1, "1H") {}
    explicit SolidSys(int nspins) : names_(nspins, "1H") {}

    // SolidSys::read: numspin/spins, T <iso> <spin>, C, D, J lines (sys.cc:120-204,
    // csa.cc:400-440, dip.cc:285-323, j.cc:171-202).
    void read(const std::vector<std::string> &lines) {
        std::vector<std::vector<std::string>> toks;
        for (auto &l : lines) {
            std::istringstream is(l);
            std::vector<std::string> t;
            std::string w;
            while (is >> w) t.push_back(w);
            if (!t.empty() && t[0][0] != '#') toks.push_back(t);
        }
        int ns = 1;
        std::map<int, std::string> named;
        for (auto &t : toks) {
            if ((t[0] == "numspin" || t[0] == "spins") && t.size() >= 2) ns = std::atoi(t[1].c_str());
            if (t[0] == "T" || t[0] == "type" || t[0] == "Type") {
                if (t.size() < 3) throw BL_exception("'T' specifier with the wrong number of inputs");
                named[std::atoi(t[2].c_str())] = t[1];
            }
        }
        if (ns < 1) throw BL_exception("numspin must be >= 1");
        names_.assign(ns, "1H");
        for (auto &kv : named) {
            if (kv.first < 0 || kv.first >= ns) throw BL_exception("spin index out of range");
            names_[kv.first] = kv.second;
        }
        csa_.clear();
        dip_.clear();
        jc_.clear();
        for (auto &t : toks) {
            auto d = [&](size_t i) { return i < t.size() ? std::atof(t[i].c_str()) : 0.0; };
            if (t[0] == "C" || t[0] == "CSA" || t[0] == "csa") {
                if (t.size() <= 4)
                    throw BL_exception(" A CSA must have at least 4 parameters.. C <iso> <del> "
                                       "<eta> <spin> <alpha> <beta> <gamma>");
                if (d(3) > 1.0 || d(3) < 0.0)
                    throw BL_exception(" 'eta' set to a negative or value >1...");
                double v[7] = {d(1), d(2), d(3), (double)std::atoi(t[4].c_str()), d(5), d(6), d(7)};
                csa_.insert(csa_.end(), v, v + 7);
            } else if (t[0] == "D" || t[0] == "Dip" || t[0] == "dip" || t[0] == "DIP") {
                if (t.size() < 4)
                    throw BL_exception(" A dipole must have at least 4 parameters..");
                double v[6] = {d(1), (double)(int)d(2), (double)(int)d(3), d(4), d(5), d(6)};
                dip_.insert(dip_.end(), v, v + 6);
            } else if (t[0] == "J" || t[0] == "j" || t[0] == "jcop" || t[0] == "Jcop") {
                if (t.size() <= 3)
                    throw BL_exception(" A j-coupling must have at least 3 parameters..");
                double v[4] = {d(1), (double)std::atoi(t[2].c_str()),
                               (double)std::atoi(t[3].c_str()),
                               t.size() >= 5 ? (double)std::atoi(t[4].c_str()) : 0.0};
                jc_.insert(jc_.end(), v, v + 4);
            }
        }
    }
    static SolidSys fromText(const std::string &txt) {
        std::vector<std::string> lines;
        std::istringstream is(txt);
        std::string l;
        while (std::getline(is, l)) lines.push_back(l);
        SolidSys s;
        s.read(lines);
        return s;
    }

    int spins() const { return (int)names_.size(); }
    int HS() const { return 1 << spins(); }
    const std::string &isotope(int i) const { return names_.at(i); }

    // single-spin operators expanded over the spin system (multispinop.cc:51-70)
    matrix Ix(int i) const { return single('x', i); }
    matrix Iy(int i) const { return single('y', i); }
    matrix Iz(int i) const { return single('z', i); }
    matrix Ip(int i) const { return single('p', i); }
    matrix Im(int i) const { return single('m', i); }
    matrix Fx() const { return total('x'); }
    matrix Fy() const { return total('y'); }
    matrix Fz() const { return total('z'); }
    matrix Fp() const { return total('p'); }
    matrix Fm() const { return total('m'); }
    matrix Fe() const { return matrix::identity(HS()); }

    // Operator strings: sums of [coef*]I{x,y,z,p,m,e}[_spin] (HamiltonianGen subset)
    matrix op(const std::string &expr) const {
        matrix out(HS());
        std::string e;
        for (char ch : expr)
            if (ch != ' ') e += ch;
        size_t i = 0;
        while (i < e.size()) {
            double sign = 1.0;
            if (e[i] == '+' || e[i] == '-') {
                sign = e[i] == '-' ? -1.0 : 1.0;
                i++;
            }
            size_t j = i;
            while (j < e.size() && !((e[j] == '+' || e[j] == '-') && j > i && e[j - 1] != '*' &&
                                     e[j - 1] != 'e' && e[j - 1] != 'E'))
                j++;
            std::string term = e.substr(i, j - i);
            matrix t = sign * matrix::identity(HS());
            size_t p = 0;
            while (p <= term.size()) {
                size_t q = term.find('*', p);
                if (q == std::string::npos) q = term.size();
                std::string f = term.substr(p, q - p);
                if (f.size() >= 2 && f[0] == 'I') {
                    char c = f[1];
                    if (f.size() == 2) t = t * total(c);
                    else if (f.size() >= 4 && f[2] == '_') t = t * single(c, std::atoi(f.c_str() + 3));
                    else throw BL_exception("cannot evaluate operator string '" + f + "'");
                } else {
                    char *endp = nullptr;
                    double v = std::strtod(f.c_str(), &endp);
                    if (f.empty() || *endp) throw BL_exception("cannot evaluate '" + f + "'");
                    t *= v;
                }
                p = q + 1;
            }
            out += t;
            i = j;
        }
        return out;
    }

    // label of an isotope in abi() tables, -1 when no spin has it
    int isotopeId(const std::string &name) const {
        std::map<std::string, int> ids;
        for (int i = 0; i < spins(); i++)
            if (ids.find(names_[i]) == ids.end()) ids.insert({names_[i], (int)ids.size()});
        auto it = ids.find(name);
        return it == ids.end() ? -1 : it->second;
    }

    // flat tables for the C ABI
    blfid_system abi() const {
        iso_ids_.assign(spins(), 0);
        std::map<std::string, int> ids;
        for (int i = 0; i < spins(); i++) {
            auto it = ids.find(names_[i]);
            if (it == ids.end()) it = ids.insert({names_[i], (int)ids.size()}).first;
            iso_ids_[i] = it->second;
        }
        blfid_system s;
        s.nspins = spins();
        s.isotope = iso_ids_.data();
        s.ncsa = (int)(csa_.size() / 7);
        s.csa = csa_.data();
        s.ndip = (int)(dip_.size() / 6);
        s.dip = dip_.data();
        s.nj = (int)(jc_.size() / 4);
        s.jc = jc_.data();
        return s;
    }

  private:
    matrix single(char c, int spin) const {
        if (spin < 0 || spin >= spins()) throw BL_exception("spin index out of range");
        const int S = spins(), N = HS();
        matrix m(N);
        for (int col = 0; col < N; col++) {
            int b = (col >> (S - 1 - spin)) & 1; // 0: m=+1/2, 1: m=-1/2
            int flip = col ^ (1 << (S - 1 - spin));
            switch (c) {
            case 'z': m(col, col) = b ? -0.5 : 0.5; break;
            case 'x': m(flip, col) = 0.5; break;
            case 'y': m(flip, col) = b ? complex(0, -0.5) : complex(0, 0.5); break;
            case 'p': if (b) m(flip, col) = 1.0; break;
            case 'm': if (!b) m(flip, col) = 1.0; break;
            case 'e': m(col, col) = 1.0; break;
            default: throw BL_exception(std::string("unknown spin operator I") + c);
            }
        }
        return m;
    }
    matrix total(char c) const {
        if (c == 'e') return Fe();
        matrix m(HS());
        for (int i = 0; i < spins(); i++) m += single(c, i);
        return m;
    }

    std::vector<std::string> names_;
    std::vector<double> csa_, dip_, jc_;
    mutable std::vector<int> iso_ids_;
};

// Powder orientation set (powder.cc:333-389): rows "phi theta [gamma] weight".
class powder {
  public:
    powder() = default;
    explicit powder(const std::string &file) { read(file); }
    void read(const std::string &file) {
        std::ifstream in(file);
        if (!in) throw BL_exception("no POWDER file name like " + file);
        th_.clear(); ph_.clear(); ga_.clear(); w_.clear();
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream is(line);
            std::vector<double> v;
            std::string tok;
            while (is >> tok) {
                if (v.empty() && (tok[0] == '#' || tok[0] == '%')) break;
                v.push_back(std::atof(tok.c_str()));
            }
            if (v.empty()) continue;
            ph_.push_back(v[0]);
            th_.push_back(v.size() >= 2 ? v[1] : 0.0);
            if (v.size() == 3) { w_.push_back(v[2]); ga_.push_back(0.0); }
            else if (v.size() >= 4) { ga_.push_back(v[2]); w_.push_back(v[3]); }
            else { w_.push_back(1.0); ga_.push_back(0.0); }
        }
    }
    void add(double theta, double phi, double gamma, double w) {
        th_.push_back(theta); ph_.push_back(phi); ga_.push_back(gamma); w_.push_back(w);
    }
    int size() const { return (int)th_.size(); }
    double theta(int k) const { return th_[k]; }
    double phi(int k) const { return ph_[k]; }
    double gamma(int k) const { return ga_[k]; }
    double weight(int k) const { return w_[k]; }
    const double *thetas() const { return th_.data(); }
    const double *phis() const { return ph_.data(); }
    const double *gammas() const { return ga_.data(); }
    const double *weights() const { return w_.data(); }

  private:
    std::vector<double> th_, ph_, ga_, w_;
};

// oneFID<SolidSys>: the GPU engine behind BlochLib's driver API.  FID(powder&, ro, det)
// returns sum_k w_k fid_k (oneFID.h:396-402, no Tr(DD^H) division); solid-2.0's doFID
// divides by trace(det, adjoint(det)) itself (sequencerun_funcs.cc:698-702).
template <class Func_T = SolidSys> class oneFID {
  public:
    oneFID() : function_(nullptr), npts_(0), sw_(20000), wr_(0), rotb_(0), startT_(0) {}
    oneFID(Func_T &in, int npts, double sw = 20000, double wr = 0, double rotb = 0,
           double startT = 0)
        : function_(&in), npts_(npts), sw_(sw), wr_(wr), rotb_(rotb), startT_(startT) {}

    void setFunction(Func_T &in) { function_ = &in; }
    void setFunction(Func_T *in) { function_ = in; }
    void setWr(double v) { wr_ = v; }
    void setRotorSpeed(double v) { wr_ = v; }
    void setRotorAngle(double v) { rotb_ = v; }
    void setRotorBeta(double v) { rotb_ = v; }
    void setSw(double v) { sw_ = v; }
    void setSweepWidth(double v) { sw_ = v; }
    void setSize(int n) { npts_ = n; }
    void setNpts(int n) { npts_ = n; }
    void npts(int n) { npts_ = n; }
    void setStartTime(double t) { startT_ = t; }
    double wr() const { return wr_; }
    double rotorAngle() const { return rotb_; }
    double sw() const { return sw_; }
    int npts() const { return npts_; }
    int size() const { return npts_; }
    double startTime() const { return startT_; }

    std::vector<complex> FID(const powder &pw, const matrix &ro, const matrix &det) {
        return run(pw.size(), pw.thetas(), pw.phis(), pw.gammas(), pw.weights(), ro, det);
    }
    std::vector<complex> FID(const matrix &ro, const matrix &det, double the = 0, double phi = 0,
                             double gam = 0) {
        double w = 1.0;
        return run(1, &the, &phi, &gam, &w, ro, det);
    }

  private:
    std::vector<complex> run(int K, const double *th, const double *ph, const double *ga,
                             const double *w, const matrix &ro, const matrix &det) {
        if (!function_) throw BL_exception("oneFID::FID(...) NO Function defined");
        if (npts_ <= 0)
            throw BL_exception("Fid size was '0' please use 'npts(int)' set set it");
        blfid_system s = function_->abi();
        blfid_plan *plan = nullptr;
        check(blfid_plan_create(&s, wr_, rotb_, sw_, npts_, startT_, BLFID_FP64, &plan));
        std::vector<complex> out(npts_);
        blfid_status st = blfid_powder_fid(plan, th, ph, ga, w, K, ro.data(), det.data(), 0,
                                           reinterpret_cast<double *>(out.data()));
        double swa = blfid_plan_sw(plan);
        blfid_plan_destroy(plan);
        check(st);
        sw_ = swa;
        return out;
    }

    Func_T *function_;
    int npts_;
    double sw_, wr_, rotb_, startT_;
};

// rf over one rotor period of `steps` equal steps: one channel per pulsed isotope.  pulse()
// takes the script's units: amplitude and offset in Hz, phase in degrees (converted to radians
// as pulsedata.cc:213-215 does).  A step where a channel's amplitude is 0 adds nothing for it.
class pulseSequence {
  public:
    explicit pulseSequence(int steps) : steps_(steps) {
        if (steps < 1) throw BL_exception("a pulse sequence needs at least one step");
    }
    int steps() const { return steps_; }
    // steps [k0, k1) of isotope `on` get amp (cos(phase) Ix + sin(phase) Iy) + offset Iz
    void pulse(const std::string &on, int k0, int k1, double amp, double phase_deg,
               double offset = 0.0) {
        if (k0 < 0 || k1 > steps_ || k0 > k1) throw BL_exception("pulse steps out of range");
        channel &c = chan(on);
        for (int k = k0; k < k1; k++) {
            c.amp[k] = amp;
            c.phase[k] = phase_deg * 0.017453292519943295769236907684886;
            c.offset[k] = offset;
        }
    }
    // list of blfid_rf for a system (channels on isotopes absent from it map to label -1)
    template <class Func_T> std::vector<blfid_rf> abi(const Func_T &sys) const {
        std::vector<blfid_rf> out;
        for (auto &c : chans_)
            out.push_back({sys.isotopeId(c.on), c.amp.data(), c.phase.data(), c.offset.data()});
        return out;
    }

  private:
    struct channel {
        std::string on;
        std::vector<double> amp, phase, offset;
    };
    channel &chan(const std::string &on) {
        for (auto &c : chans_)
            if (c.on == on) return c;
        chans_.push_back({on, std::vector<double>(steps_, 0.0), std::vector<double>(steps_, 0.0),
                          std::vector<double>(steps_, 0.0)});
        return chans_.back();
    }
    int steps_;
    std::vector<channel> chans_;
};

// Rotor-synchronised point-to-point acquisition of a one-period pulse program: point m is
// sum_k w_k Tr(rho_m det) [/ Tr(det det^H) when normalize, as doFID does], rho_{m+1} =
// U rho_m U^H, U = prod_k exp(+i 2 pi tau (H_k + H_rf,k)).
template <class Func_T = SolidSys> class ptopFID {
  public:
    ptopFID(Func_T &sys, int npts, double wr, double rotb, const pulseSequence &seq,
            double startT = 0)
        : function_(&sys), npts_(npts), wr_(wr), rotb_(rotb), startT_(startT), seq_(seq) {}

    std::vector<complex> FID(const powder &pw, const matrix &ro, const matrix &det,
                             bool normalize = true) {
        if (npts_ <= 0)
            throw BL_exception("Fid size was '0' please use 'npts(int)' set set it");
        blfid_system s = function_->abi();
        std::vector<blfid_rf> rf = seq_.abi(*function_);
        blfid_plan *plan = nullptr;
        check(blfid_plan_create_ptop(&s, wr_, rotb_, seq_.steps(), npts_, startT_, (int)rf.size(),
                                     rf.data(), BLFID_FP64, &plan));
        std::vector<complex> out(npts_);
        blfid_status st = blfid_powder_fid(plan, pw.thetas(), pw.phis(), pw.gammas(),
                                           pw.weights(), pw.size(), ro.data(), det.data(),
                                           normalize ? 1 : 0,
                                           reinterpret_cast<double *>(out.data()));
        blfid_plan_destroy(plan);
        check(st);
        return out;
    }

  private:
    Func_T *function_;
    int npts_;
    double wr_, rotb_, startT_;
    pulseSequence seq_;
};

} // namespace blfid
