// opcount.cpp -- op-counting build of the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Compiles wsr_oracle.c unchanged as C++ with its
// `double` replaced by `Counted`, a scalar that tallies every arithmetic operation
// it takes part in, and with -finstrument-functions so that each operation is
// attributed to the oracle functions active when it ran.  The tallies give the
// operation count of every building block of the reference algorithm (one Roe
// average, one characteristic projection, one WENO5-JS / -Z / -SR scalar
// reconstruction and its parts, one HLLC flux, ...), which is the numerator of the
// benchmark's roofline.
//
// Counting rules: add and subtract, multiply, divide, sqrt, exp, log (and log10)
// and tanh count one each; comparisons, min / max (written as comparisons in the
// oracle), fabs and sign flips count zero; expressions of literals only are
// constants folded at compile time and count zero.  FP32 arithmetic is not
// counted: the FP64 closure (scheme SR64) has the same operations as the FP32 one.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NI __attribute__((no_instrument_function))

namespace opc {
enum Kind { ADD = 0, MUL, DIV, SQRT, EXP, LOG, TANH, NKIND };
constexpr int MAXF = 64;
constexpr int MAXDEPTH = 64;
struct Tally {
    unsigned long long calls;
    unsigned long long incl[NKIND];
    unsigned long long excl[NKIND];
};
static const void *fn_addr[MAXF];
static const char *fn_name[MAXF];
static int nfn = 0;
static Tally tally[MAXF];
static int stack[MAXDEPTH];
static int depth = 0;
static unsigned long long total[NKIND];

NI static inline void count(Kind k) {
    ++total[k];
    for (int d = 0; d < depth; ++d) ++tally[stack[d]].incl[k];
    if (depth) ++tally[stack[depth - 1]].excl[k];
}

NI static inline int lookup(const void *fn) {
    for (int i = 0; i < nfn; ++i)
        if (fn_addr[i] == fn) return i;
    return -1;
}
}  // namespace opc

struct Counted {
    double v;
    Counted() = default;
    NI Counted(double x) : v(x) {}
    NI explicit operator double() const { return v; }
    NI explicit operator float() const { return (float)v; }
    NI explicit operator int() const { return (int)v; }
    NI explicit operator long() const { return (long)v; }
    NI explicit operator long long() const { return (long long)v; }
    NI explicit operator unsigned long() const { return (unsigned long)v; }
    NI explicit operator bool() const { return v != 0.0; }
    NI Counted &operator+=(Counted b) { opc::count(opc::ADD); v += b.v; return *this; }
    NI Counted &operator-=(Counted b) { opc::count(opc::ADD); v -= b.v; return *this; }
    NI Counted &operator*=(Counted b) { opc::count(opc::MUL); v *= b.v; return *this; }
    NI Counted &operator/=(Counted b) { opc::count(opc::DIV); v /= b.v; return *this; }
};
static_assert(sizeof(Counted) == sizeof(double), "Counted must have the layout of a double");

NI static inline Counted operator+(Counted a, Counted b) { opc::count(opc::ADD); return Counted(a.v + b.v); }
NI static inline Counted operator-(Counted a, Counted b) { opc::count(opc::ADD); return Counted(a.v - b.v); }
NI static inline Counted operator*(Counted a, Counted b) { opc::count(opc::MUL); return Counted(a.v * b.v); }
NI static inline Counted operator/(Counted a, Counted b) { opc::count(opc::DIV); return Counted(a.v / b.v); }
NI static inline Counted operator-(Counted a) { return Counted(-a.v); }
NI static inline Counted operator+(Counted a) { return a; }
NI static inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
NI static inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
NI static inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
NI static inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
NI static inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
NI static inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
NI static inline Counted sqrt(Counted a) { opc::count(opc::SQRT); return Counted(::sqrt(a.v)); }
NI static inline Counted exp(Counted a) { opc::count(opc::EXP); return Counted(::exp(a.v)); }
NI static inline Counted log(Counted a) { opc::count(opc::LOG); return Counted(::log(a.v)); }
NI static inline Counted log10(Counted a) { opc::count(opc::LOG); return Counted(::log10(a.v)); }
NI static inline Counted tanh(Counted a) { opc::count(opc::TANH); return Counted(::tanh(a.v)); }
NI static inline Counted fabs(Counted a) { return Counted(::fabs(a.v)); }
NI static inline bool isfinite(Counted a) { return std::isfinite(a.v); }

#define double Counted
#include "wsr_oracle.c"
#undef double

#define TRACK(fn) {(const void *)&(fn), #fn}
static const struct { const void *addr; const char *name; } TRACKED[] = {
    TRACK(con_to_pri), TRACK(pri_to_flux), TRACK(hllc_flux), TRACK(force_flux),
    TRACK(roe_average_state), TRACK(jacobian_state_values), TRACK(con_to_char_x), TRACK(con_to_char_y),
    TRACK(char_to_con_x), TRACK(char_to_con_y), TRACK(candidates), TRACK(optimal_weights), TRACK(js_beta),
    TRACK(js_combine), TRACK(z_combine), TRACK(weno5_js_value), TRACK(weno5_z_value), TRACK(raw_sensors),
    TRACK(stencil_scale), TRACK(nn_features), TRACK(plateau_detected), TRACK(mlp_ratios_f64),
    TRACK(raw_ratios), TRACK(sr_ratios), TRACK(sr_weights), TRACK(mlp_weights), TRACK(weno5_sr_value),
    TRACK(scalar_value), TRACK(char_face), TRACK(normal_stage), TRACK(flux_stage), TRACK(update_rk3),
    TRACK(copy_boundary), TRACK(apply_double_mach), TRACK(oracle_euler2d_stage),
    TRACK(oracle_euler2d_max_speed), TRACK(oracle_euler2d_dt),
    TRACK(w7_candidates), TRACK(w7_classical_beta), TRACK(w7_features), TRACK(w7_plateau), TRACK(w7_mlp_raw64),
    TRACK(w7_sym_ratios), TRACK(w7_scalar_2d), TRACK(w7_guard_trips), TRACK(w7_optimal), TRACK(w7_beta_poly),
    TRACK(w7_beta_sos), TRACK(w7_provider_beta),
};
#undef TRACK

extern "C" {

NI void __cyg_profile_func_enter(void *fn, void *) {
    int id = opc::lookup(fn);
    if (id < 0) return;
    if (opc::depth >= opc::MAXDEPTH) abort();
    ++opc::tally[id].calls;
    opc::stack[opc::depth++] = id;
}

NI void __cyg_profile_func_exit(void *fn, void *) {
    int id = opc::lookup(fn);
    if (id < 0) return;
    if (opc::depth <= 0 || opc::stack[opc::depth - 1] != id) abort();
    --opc::depth;
}

NI int opcount_init(void) {
    opc::nfn = 0;
    for (const auto &t : TRACKED) {
        if (opc::nfn >= opc::MAXF) return -1;
        opc::fn_addr[opc::nfn] = t.addr;
        opc::fn_name[opc::nfn] = t.name;
        ++opc::nfn;
    }
    memset(opc::tally, 0, sizeof opc::tally);
    memset(opc::total, 0, sizeof opc::total);
    opc::depth = 0;
    return opc::nfn;
}

NI int opcount_nkinds(void) { return opc::NKIND; }
NI int opcount_nfunctions(void) { return opc::nfn; }
NI const char *opcount_name(int i) { return (i >= 0 && i < opc::nfn) ? opc::fn_name[i] : ""; }

/* out: calls, incl[NKIND], excl[NKIND] */
NI int opcount_tally(int i, unsigned long long *out) {
    if (i < 0 || i >= opc::nfn) return -1;
    out[0] = opc::tally[i].calls;
    for (int k = 0; k < opc::NKIND; ++k) out[1 + k] = opc::tally[i].incl[k];
    for (int k = 0; k < opc::NKIND; ++k) out[1 + opc::NKIND + k] = opc::tally[i].excl[k];
    return 0;
}

NI void opcount_total(unsigned long long *out) {
    for (int k = 0; k < opc::NKIND; ++k) out[k] = opc::total[k];
}

static Counted *to_counted(const double *src, size_t n) {
    Counted *dst = (Counted *)malloc(sizeof(Counted) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) dst[i] = Counted(src[i]);
    return dst;
}

/* One counted oracle stage (oracle_euler2d_stage) on the padded FP64 state u
 * (ghost 3); mlp is the 189 FP64 parameters for scheme SR64, else NULL. */
NI int opcount_euler2d_stage(const double *u, int nx, int ny, double dx, double dy, double gamma, double dt,
                             int rk, int scheme, int riemann, int boundary, double boundary_time, int eno_cutoff,
                             const double *mlp) {
    if (scheme != SCHEME_JS && scheme != SCHEME_Z && scheme != SCHEME_SR64) return -2;
    size_t n = (size_t)(nx + 6) * (size_t)(ny + 6) * 4u;
    Counted *cu = to_counted(u, n);
    Counted *out = to_counted(u, n);
    Counted *work = (Counted *)malloc(sizeof(Counted) * oracle_euler2d_work_doubles(nx, ny, 3));
    Counted *cm = mlp ? to_counted(mlp, MLP5_COUNT) : NULL;
    int rc = oracle_euler2d_stage(cu, cu, out, nx, ny, 3, Counted(dx), Counted(dy), Counted(gamma), Counted(dt), rk,
                                  scheme, riemann, boundary, Counted(boundary_time), eno_cutoff, cm, work);
    free(cu);
    free(out);
    free(work);
    free(cm);
    return rc;
}

/* One counted Roe average of two conserved states. */
NI void opcount_roe(const double *l, const double *r, double gamma, double *out) {
    Counted a[4], b[4], o[4];
    for (int k = 0; k < 4; ++k) { a[k] = Counted(l[k]); b[k] = Counted(r[k]); }
    oracle_roe(a, b, Counted(gamma), o);
    for (int k = 0; k < 4; ++k) out[k] = o[k].v;
}

/* One counted projection onto the characteristic fields of roe and back. */
NI void opcount_char_roundtrip(const double *q, const double *roe, int direction, double gamma, double *c,
                               double *back) {
    Counted a[4], b[4], oc[4], ob[4];
    for (int k = 0; k < 4; ++k) { a[k] = Counted(q[k]); b[k] = Counted(roe[k]); }
    oracle_char_roundtrip(a, b, direction, Counted(gamma), oc, ob);
    for (int k = 0; k < 4; ++k) { c[k] = oc[k].v; back[k] = ob[k].v; }
}

/* One counted scalar reconstruction of the 5-value stencil q for one head. */
NI double opcount_scalar(const double *q, int head, int scheme, const double *mlp, int eno_cutoff, double h) {
    Counted a[5];
    for (int k = 0; k < 5; ++k) a[k] = Counted(q[k]);
    Counted *cm = mlp ? to_counted(mlp, MLP5_COUNT) : NULL;
    double v = oracle_scalar_value(a, head, scheme, cm, eno_cutoff, Counted(h)).v;
    free(cm);
    return v;
}

/* One counted WENO7 scalar reconstruction of the 7-value stencil q (FP64 closure
 * parameters, 6-24-16-16-4, for scheme SR64). */
NI double opcount_scalar7(const double *q, int gauss, int lr, int scheme, const double *mlp, int eno, double h) {
    Counted a[7];
    for (int k = 0; k < 7; ++k) a[k] = Counted(q[k]);
    Counted *cm = mlp ? to_counted(mlp, MLP7_COUNT) : NULL;
    int trip = 0;   /* with a trip flag the Z-fallback guard runs (eno bit 1) */
    double v = w7_scalar_2d(a, gauss, lr, scheme, cm, eno, Counted(h), &trip).v;
    free(cm);
    return v;
}

/* One counted HLLC flux between two conserved states. */
NI void opcount_hllc(const double *ul, const double *ur, int direction, double gamma, double *out) {
    Counted l[4], r[4], o[4];
    for (int k = 0; k < 4; ++k) { l[k] = Counted(ul[k]); r[k] = Counted(ur[k]); }
    oracle_hllc(l, r, direction, Counted(gamma), o);
    for (int k = 0; k < 4; ++k) out[k] = o[k].v;
}

}  // extern "C"
