// hhg_device.hpp — device-side building blocks of the Bloch hot path (gfx950).
//
// One wavefront lane owns one k-point.  Everything is FP64.  The functions
// here are the MI355X formulation of:
//   Laser::laser_function            sources/HHG/Laser/Laser.hpp:32-54
//   CosineLaser / GaussLaser envelope sources/HHG/Laser/CosineLaser.cpp:22-26,
//                                     sources/HHG/Laser/GaussLaser.cpp:26-28
//   DoubleLaser A + B(t - delay)      sources/HHG/Laser/DoubleLaser.hpp:63-67
//   Bloch RHS ds/dt = 2 d x s         sources/HHG/Systems/PiFlux.cpp:89-94,
//                                     sources/HHG/Systems/Honeycomb.cpp:130-138
//   relaxation to the initial state   sources/HHG/Systems/DiracSystem.cpp:231-233
//   odeint RKF78 controller           Boost.Odeint controlled_runge_kutta (call
//                                     sites sources/HHG/Systems/DiracSystem.cpp:167-173)
//   current vertices                  sources/HHG/Systems/PiFlux.cpp:831-838,
//                                     sources/HHG/Systems/Honeycomb.cpp:273-279
#pragma once
#include <hip/hip_runtime.h>

#include "hhg_internal.hpp"

namespace hhg {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kSqrt3 = 1.73205080756887729352744634150587237;
constexpr double kHalfSqrt3 = 0.866025403784438646763723170752936183;
constexpr double kDblEps = 2.220446049250313080847263336181640625e-16;  // odeint's less_with_sign epsilon

// Job records are read-only during a launch; referencing them through the
// constant address space lets every field be fetched with a scalar load when
// it is used instead of being held in registers.
using ConstJob = const __attribute__((address_space(4))) JobDev;
using JobRef = ConstJob&;

__device__ __forceinline__ JobRef job_ref(const JobDev* p) {
    return *reinterpret_cast<ConstJob*>(reinterpret_cast<uintptr_t>(p));
}

// The same record through an opaque copy of its address.  A field read
// through the copy cannot be hoisted above the point that makes it, so a loop
// that makes the copy in every pass fetches the field with a scalar load where
// it is used; read through the original, every field the loop needs is loaded
// once before it and held in scalar registers for the whole loop, and the ones
// that do not fit are spilled to VGPR lanes (a VALU slot at every use).
__device__ __forceinline__ JobRef fresh_job(JobRef J) {
    uint64_t p = reinterpret_cast<uint64_t>(&J);
    asm volatile("" : "+s"(p));
    return *reinterpret_cast<ConstJob*>(p);
}

// ---------------------------------------------------------------------------
// Pulses: A(t) = amp * envelope * sin(t + cep); B = same shape at t - delay.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double env_cos2(JobRef J, double t) {
    if (t <= J.t0 || t >= J.t_env_end) return 0.0;
    return 0.5 * (1.0 - cos((t - J.t0) * J.inv_n));
}

__device__ __forceinline__ double env_gauss(JobRef J, double t) {
    const double x = t - J.t0;
    return exp(-(x * x) * J.inv_2s2);
}

// Spline pulse: the clamped cubic spline of the samples, per interval a cubic
// in x = (t - t_i) / h (coefficients built on the host); the lanes of a wave
// usually evaluate the same interval, so the 32-byte row is one cached load.
__device__ __forceinline__ double spline_A(JobRef J, double t) {
    if (!(t >= J.spl_t0) || t > J.spl_t_last) return 0.0;
    const double u = (t - J.spl_t0) * J.spl_inv_h;
    int i = (int)u;
    i = i < J.spl_n - 1 ? i : J.spl_n - 1;
    const double x = u - (double)i;
    const double* c = J.spl + 4 * (int64_t)i;
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}

template <int P>
__device__ __forceinline__ double pulse_A(JobRef J, double t) {
    if constexpr (P == PULSE_SPLINE) {
        return spline_A(J, t);
    } else if constexpr (P == PULSE_COS2) {
        return J.amp_a * env_cos2(J, t) * sin(t + J.cep);
    } else if constexpr (P == PULSE_GAUSS) {
        return J.amp_a * env_gauss(J, t) * sin(t + J.cep);
    } else if constexpr (P == PULSE_DOUBLE_GAUSS) {
        const double tb = t - J.delay;
        return J.amp_a * env_gauss(J, t) * sin(t + J.cep) + J.amp_b * env_gauss(J, tb) * sin(tb + J.cep);
    } else if constexpr (P == PULSE_DOUBLE_COS2) {
        const double tb = t - J.delay;
        return J.amp_a * env_cos2(J, t) * sin(t + J.cep) + J.amp_b * env_cos2(J, tb) * sin(tb + J.cep);
    } else {
        return J.amp_a * sin(t + J.cep);  // continuous: envelope 1
    }
}

// ---------------------------------------------------------------------------
// Per-step pulse cache.  At the start of a step (time t) the carrier and the
// cos^2 envelope phase are evaluated once with full-range sincos; the stage
// times t + c_i h then follow by exact angle addition with a small-argument
// series.  The series needs |c_i h| <= 0.25 and |c_i h / n| <= 0.25; steps
// never exceed the sampling interval, and the plan rejects grids with
// Delta > 0.25, so the bound always holds.  Gaussian envelopes are evaluated
// directly with exp at every stage.
// ---------------------------------------------------------------------------
constexpr double kSmallArg = 0.25;

// sin x and cos x for |x| <= 0.25 (Taylor; remainder < 3e-20 relative).
__device__ __forceinline__ void small_sincos(double x, double& sn, double& cs) {
    const double x2 = x * x;
    sn = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
                                                                                          x2 * (-1.0 / 39916800.0))))));
    cs = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 + x2 * (-1.0 / 3628800.0 +
                                                                                                x2 * (1.0 / 479001600.0))))));
}

// One pulse component (A or B) at a step start.
struct PulseComp {
    double S, C;    // sin, cos of the carrier phase
    double E0, E1;  // cos^2 only: sin, cos of the envelope phase
};

template <int P>
struct PulseCache {
    static constexpr bool kGauss = (P == PULSE_GAUSS || P == PULSE_DOUBLE_GAUSS);
    static constexpr bool kDouble = (P == PULSE_DOUBLE_GAUSS || P == PULSE_DOUBLE_COS2);
    static constexpr bool kCont = (P == PULSE_CONTINUOUS);
    static constexpr bool kSpline = (P == PULSE_SPLINE);
    double t;
    PulseComp a, b;

    __device__ __forceinline__ void prep_comp(JobRef J, double tc, PulseComp& c) {
        sincos(tc + J.cep, &c.S, &c.C);
        if constexpr (!kGauss && !kCont) sincos((tc - J.t0) * J.inv_n, &c.E0, &c.E1);
    }

    __device__ __forceinline__ void prepare(JobRef J, double t_) {
        t = t_;
        if constexpr (kSpline) return;
        prep_comp(J, t, a);
        if constexpr (kDouble) prep_comp(J, t - J.delay, b);
    }

    // envelope * carrier of one component at tc + dlt, with (sd, cd) =
    // (sin dlt, cos dlt)
    __device__ __forceinline__ double comp_at(JobRef J, const PulseComp& c, double tc, double dlt, double sd,
                                              double cd) const {
        const double carrier = c.S * cd + c.C * sd;
        const double tt = tc + dlt;
        if constexpr (kCont) {
            return carrier;
        } else if constexpr (kGauss) {
            const double x = tt - J.t0;
            return exp(-(x * x) * J.inv_2s2) * carrier;
        } else {
            double sn, cs;
            small_sincos(dlt * J.inv_n, sn, cs);
            const double ce = c.E1 * cs - c.E0 * sn;  // cos(envelope phase + dlt / n)
            const double v = 0.5 * (1.0 - ce) * carrier;
            return (tt <= J.t0 || tt >= J.t_env_end) ? 0.0 : v;
        }
    }

    // A(t + dlt)
    __device__ __forceinline__ double at(JobRef J, double dlt) const {
        if constexpr (kSpline) return spline_A(J, t + dlt);
        double sd, cd;
        small_sincos(dlt, sd, cd);
        double r = J.amp_a * comp_at(J, a, t, dlt, sd, cd);
        if constexpr (kDouble) r += J.amp_b * comp_at(J, b, t - J.delay, dlt, sd, cd);
        return r;
    }
};

// ---------------------------------------------------------------------------
// Models.  Honeycomb g(q) = e^{i qx} + 2 e^{-i qx/2} cos(sqrt3/2 qy) is the
// nearest-neighbour sum of Honeycomb.cpp:44-50 with the two oblique bonds
// combined, so one sincos(qx/2) replaces three complex exponentials; with the
// field along x (theta = 0) the cos(sqrt3/2 ky) factor is a per-lane constant.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double fermi(double e, double beta) {
    if (!(beta > 0.0)) return e != 0.0 ? (e < 0.0 ? 1.0 : 0.0) : 0.5;
    return 1.0 / (1.0 + exp(beta * e));
}

// is_zero of the reference (GlobalDefinitions.hpp:61-67): biased exponent
// below 0b01111010011, i.e. |x| < 2^-44.
__device__ __forceinline__ bool ref_is_zero(double x) { return fabs(x) < 0x1p-44; }
constexpr double kRescue = 1e-4;  // RESCUE_TRAFO, PiFlux.cpp:50

// Lattice models take the field along x (theta = 0), as the reference's
// Honeycomb (Honeycomb.hpp:88-90); the Dirac model accepts any theta; the 3D
// models take it along z.  STICKY (3D pi-flux, equilibrium relaxation): the
// shifted momentum of the equilibrium targets persists across right-hand-side
// calls, as in PiFlux::evolve_occupation_numbers (PiFlux.cpp:238, 252-257), so
// its rescued cos x stays once set.
template <int M, int P, int R, bool STICKY = false>
struct Bloch {
    static constexpr int kModel = M;
    static constexpr int kRelax = R;
    // stage values are (sin f A, cos f A); the occupation numbers' integration
    // (STICKY) takes them too, so that it computes cos(kz - A) with the same
    // operations as the plain relaxing trajectory it is checked against
    static constexpr bool kPhase = phase_table(M, R);
    JobRef J;
    double kx, ky, kz;
    double pk_s = 0.0, pk_c = 0.0;  // phase models: sin and cos of f kx (f kz for the 3D pi-flux lattice)
    double s0[3];
    double cy;  // theta = 0 constant: PiFlux 2t cos ky, honeycomb cos(sqrt3/2 ky); 3D Dirac |k|; 3D pi-flux cos ky
    double ik;  // 3D Dirac: 1/|k|
    double cx;  // 3D pi-flux: cos kx (the reference's rescued value where it applies)
    double eq[3];  // equilibrium relaxation: the state at the start of the step
    bool rescued;  // 3D pi-flux, uniform grid: the reference's rescue applies to this point
    mutable double eq_cx = 0.0;    // STICKY: cos x of the persistent shifted momentum
    mutable double last_cz = 0.0;  // STICKY: its cos z as left by the last right-hand-side call

    __device__ __forceinline__ Bloch(JobRef job, double kx_, double ky_, double kz_ = 0.0)
        : J(job), kx(kx_), ky(ky_), kz(kz_) {
        ik = 0.0;
        cx = 0.0;
        rescued = false;
        if constexpr (M == MODEL_PIFLUX2D) cy = 2.0 * J.hop * cos(ky);
        else if constexpr (M == MODEL_HONEYCOMB) cy = cos(kHalfSqrt3 * ky);
        else if constexpr (M == MODEL_DIRAC3D) cy = sqrt(kx * kx + ky * ky), ik = cy > 0.0 ? 1.0 / cy : 0.0;
        else if constexpr (M == MODEL_PIFLUX3D) {
            cx = cos(kx), cy = cos(ky);
            // uniform grid (PiFlux.cpp:978-986): where cos x and cos y vanish and
            // cos z < 0 the reference replaces cos x by RESCUE |cos z| for the run
            if (J.ic_rescue == 2) {
                const double c = cos(kz);
                if (ref_is_zero(cx) && ref_is_zero(cy) && c < 0.0) cx = fabs(c) * kRescue, rescued = true;
            }
        }
        else cy = 0.0;
        if constexpr (model_has_phase(M)) sincos(model_phase_scale(M) * (M == MODEL_PIFLUX3D ? kz : kx), &pk_s, &pk_c);
    }

    // phase models: sin and cos of f (k - A) from (sin f A, cos f A)
    __device__ __forceinline__ void shifted_phase(double sa, double ca, double& sn, double& cs) const {
        sn = pk_s * ca - pk_c * sa;
        cs = pk_c * ca + pk_s * sa;
    }

    // the integrator calls this with the state at the start of every attempt
    __device__ __forceinline__ void begin_step(const double s[3]) {
        if constexpr (R == RELAX_EQUILIBRIUM) {
            eq[0] = s[0];
            eq[1] = s[1];
            eq[2] = s[2];
        }
    }

    // d(k - A e_theta); the 3D Dirac cone in the reference's band frame
    // (DiracSystem.cpp:143-153; kx = k_z, ky = kappa, field along z):
    // d = -(A kappa/|k|, 0, |k| - A k_z/|k|)
    __device__ __forceinline__ void d_of(double A, double d[3]) const {
        if constexpr (M == MODEL_PIFLUX3D) {
            // d = 2t (cos kx, cos ky, cos(kz - A)) (PiFlux.cpp:89-94, prefactor 4t = 2 * 2t)
            const double t2 = 2.0 * J.hop;
            d[0] = t2 * cx;
            d[1] = t2 * cy;
            d[2] = t2 * cos(kz - A);
        } else if constexpr (M == MODEL_DIRAC3D) {
            d[0] = -A * ky * ik;
            d[1] = 0.0;
            d[2] = A * kx * ik - cy;
        } else if constexpr (M == MODEL_DIRAC2D) {
            d[0] = J.v * (kx - A * J.ct);
            d[1] = J.v * (ky - A * J.st);
            d[2] = J.mass;
        } else if constexpr (M == MODEL_PIFLUX2D) {
            d[0] = 2.0 * J.hop * cos(kx - A);
            d[1] = cy;
            d[2] = 0.0;
        } else {
            double sh, ch;
            sincos(0.5 * (kx - A), &sh, &ch);
            honeycomb_d(sh, ch, d);
        }
    }

    // honeycomb d from (sin, cos) of qx/2, qx = kx - A
    __device__ __forceinline__ void honeycomb_d(double sh, double ch, double d[3]) const {
        const double cq = (ch - sh) * (ch + sh);  // cos qx
        const double sq = 2.0 * sh * ch;          // sin qx
        d[0] = J.hop * (cq + 2.0 * ch * cy);      // t Re g
        d[1] = -J.hop * (sq - 2.0 * sh * cy);     // -t Im g
        d[2] = 0.0;
    }

    // Honeycomb::ic_sigma (Honeycomb.cpp:355-361) at the field d = t(Re g, -Im g, 0),
    // e = |d|, with alpha^2 = f(E_F - t|g|), beta^2 = f(E_F + t|g|)
    // (Honeycomb.cpp:347-353); also the honeycomb's instantaneous equilibrium.
    __device__ __forceinline__ void honeycomb_state(const double d[3], double e, double s[3]) const {
        const double g = e / J.hop;
        const double a2 = fermi(J.fermi - J.hop * g, J.beta);
        const double b2 = fermi(J.fermi + J.hop * g, J.beta);
        const double diff = a2 - b2, prod = 2.0 * sqrt(a2 * b2);
        if (e == 0.0) {
            s[0] = s[1] = 0.0;
        } else {
            s[0] = -diff * (d[0] / J.hop) / g;
            s[1] = diff * (-d[1] / J.hop) / g;
        }
        s[2] = -prod;
    }

    // Initial Bloch vector: Honeycomb::ic_sigma (Honeycomb.cpp:355-361) with
    // the occupations of Honeycomb.cpp:347-353; the 2D Dirac/PiFlux models use
    // the band-diagonal thermal state (n_up - n_low) d^, = -d^ at T = 0, E_F = 0.
    __device__ __forceinline__ void initial(double s[3]) {
        double d[3];
        // the relaxing honeycomb starts in the equilibrium at A(t_begin) (Honeycomb.cpp:193-194)
        if constexpr (M == MODEL_HONEYCOMB && R == RELAX_EQUILIBRIUM) d_of(pulse_A<P>(J, J.t_begin), d);
        else d_of(0.0, d);
        const double e = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if constexpr (M == MODEL_PIFLUX3D) {
            // PiFlux::ic_sigma (PiFlux.cpp:519-549), alpha^2 = f(-E_F + 2t|m|),
            // beta^2 = f(-E_F - 2t|m|) (PiFlux.cpp:420-426).  On the continuum
            // grid the reference scales cos kz of the positive z nodes with
            // cos kz < 0 by 1 - RESCUE^2/2 (PiFlux.cpp:825-829; the -z node of
            // PiFlux.cpp:845-849 keeps the last x/y cosines and is not rescued).
            double cz = cos(kz);
            if ((J.ic_rescue == 1 && kz > 0.0 && cz < 0.0) || rescued) cz *= (1.0 - 0.5 * kRescue * kRescue);
            const double disp = sqrt(cx * cx + cy * cy + cz * cz);
            const double a2 = fermi(-J.fermi + 2.0 * J.hop * disp, J.beta);
            const double b2 = fermi(-J.fermi - 2.0 * J.hop * disp, J.beta);
            const double diff = a2 - b2, prod = 2.0 * sqrt(a2 * b2);
            const double ze = cz + disp, nrm = disp * ze;
            s[0] = (diff * cx * ze + prod * (cy * cy + cz * ze)) / nrm;
            s[1] = (diff * cy * ze - prod * cx * cy) / nrm;
            s[2] = (diff * cz * ze - prod * cx * ze) / nrm;
        } else if constexpr (M == MODEL_DIRAC3D) {
            // DiracSystem.cpp:155-158: alpha0 = f(E_F + |k|), beta0 = f(E_F - |k|)
            const double a0 = fermi(J.fermi + cy, J.beta), b0 = fermi(J.fermi - cy, J.beta);
            s[0] = 2.0 * a0 * b0;
            s[1] = 0.0;
            s[2] = a0 * a0 - b0 * b0;
        } else if constexpr (M == MODEL_HONEYCOMB) {
            honeycomb_state(d, e, s);
        } else {
            if (e == 0.0) {
                s[0] = s[1] = s[2] = 0.0;
            } else {
                const double nu = fermi(e - J.fermi, J.beta);
                const double nl = fermi(-e - J.fermi, J.beta);
                const double f = (nu - nl) / e;
                s[0] = f * d[0];
                s[1] = f * d[1];
                s[2] = f * d[2];
            }
        }
        s0[0] = s[0];
        s0[1] = s[1];
        s0[2] = s[2];
    }

    // the pulse at the stage times of a step: pc.prepare at its start t, then
    // pc.at(dlt) = A(t + dlt)
    PulseCache<P> pc;

    // Fermi pair of the equilibrium relaxations: P = f(c + x), M = f(c - x) for
    // x >= 0, f(e) = 1 / (1 + exp(beta e)) (GlobalDefinitions.hpp:36-41).  With
    // q = exp(-beta x) in [0, 1] and w = exp(beta c) (JobDev::fermi_w),
    // P = q / (q + w) and M = 1 / (1 + w q), so
    //   pm = P - M = w (q^2 - 1) / ((q + w)(1 + w q)),  pp = P M = q / ((q + w)(1 + w q)):
    // one expm1 and one division instead of two exponentials and two divisions.
    // q^2 - 1 is formed as (q - 1)(q + 1) from q - 1 = expm1(-beta x), which keeps
    // its relative accuracy where beta x is small (P - M itself cancels there).
    // w in [1e-150, 1e150] keeps (q + w)(1 + w q) finite and nonzero; w = 0 (T = 0,
    // or |beta c| > 345) evaluates the two Fermi functions as they stand.
    __device__ __forceinline__ void fermi_pair(double c, double x, double& pm, double& pp) const {
        const double w = J.fermi_w;
        if (w > 0.0) {
            const double q = exp(-J.beta * x);
            const double r = 1.0 / ((q + w) * fma(w, q, 1.0));
            pm = w * ((q - 1.0) * (q + 1.0)) * r;
            pp = q * r;
        } else {
            const double fp = fermi(c + x, J.beta), fm = fermi(c - x, J.beta);
            pm = fp - fm;
            pp = fp * fm;
        }
    }

    // Relaxation of the 3D pi-flux lattice toward the instantaneous
    // equilibrium at m = (cos kx, cos ky, cz = cos(kz - A)): update_equilibrium_state
    // (PiFlux.cpp:134-166) with the off-diagonal phase of the step-start state
    // in the band frame (diagonal_sigma, PiFlux.cpp:1037-1049), subtracted as
    // rd (s/3 - D) + ro (2s/3 - O) (PiFlux.cpp:174-179).  alpha^2 - beta^2 and
    // alpha beta come from fermi_pair with c = -E_F, x = 2t|m|.  The band-frame
    // components g = G / nz (nz = 2|m| z >= 0) enter only through their phase
    // G / |G|, so G is not divided by nz; the reference's is_zero(|g|) test is
    // |G| < 2^-44 nz.  Three divisions, three square roots and one expm1 per
    // call (the reference's form takes eight, three and two exponentials).
    __device__ __forceinline__ void relax_equilibrium(double cz, const double s[3], double out[3]) const {
        double mx = cx;
        if constexpr (STICKY) {
            cz = shift_frame(cz);
            mx = eq_cx;
        } else if (ref_is_zero(mx) && ref_is_zero(cy) && cz < 0.0) {
            mx = fabs(cz) * kRescue;
            cz *= (1.0 - 0.5 * kRescue * kRescue);
        }
        const double my = cy;
        const double disp = sqrt(mx * mx + my * my + cz * cz);
        double diff, pp;
        fermi_pair(-J.fermi, 2.0 * J.hop * disp, diff, pp);
        const double p2 = 2.0 * sqrt(pp);
        const double z = cz + disp, nz = 2.0 * disp * z;
        const double xz = 2.0 * mx * z, yz = 2.0 * my * z, xy = 2.0 * mx * my;
        const double G0 = eq[0] * (z * z - mx * mx + my * my) - xy * eq[1] - xz * eq[2];
        const double G1 = eq[1] * (z * z + mx * mx - my * my) - xy * eq[0] - yz * eq[2];
        const double L = sqrt(G0 * G0 + G1 * G1);
        const bool flat = L < 0x1p-44 * nz;
        const double u = flat ? 0.0 : p2 / L;
        const double prod = flat ? p2 : u * G0;
        const double imag = u * G1;
        const double inv = 1.0 / (disp * z);
        const double fd = diff * z * inv;
        const double O0 = (prod * (my * my + cz * z) + imag * (-mx * my)) * inv;
        const double O1 = (prod * (-mx * my) + imag * (mx * mx + cz * z)) * inv;
        const double O2 = (prod * (-mx * z) + imag * (-my * z)) * inv;
        const double rd = J.relax_rate, ro = J.relax_rate2;
        out[0] -= rd * (s[0] * (1.0 / 3.0) - mx * fd) + ro * (s[0] * (2.0 / 3.0) - O0);
        out[1] -= rd * (s[1] * (1.0 / 3.0) - my * fd) + ro * (s[1] * (2.0 / 3.0) - O1);
        out[2] -= rd * (s[2] * (1.0 / 3.0) - cz * fd) + ro * (s[2] * (2.0 / 3.0) - O2);
    }

    // The honeycomb's instantaneous equilibrium at d (the relaxation target of
    // Honeycomb.cpp:169-177): honeycomb_state with alpha^2 - beta^2 and
    // alpha beta from fermi_pair (c = E_F, x = |d|; alpha^2 = M, beta^2 = P), so
    // (s0, s1) = (P - M) (d0, d1) / |d|.  Two divisions, two square roots and one
    // expm1 (honeycomb_state: seven, two and two exponentials).
    __device__ __forceinline__ void honeycomb_target(const double d[3], double se[3]) const {
        const double e = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double pm, pp;
        fermi_pair(J.fermi, e, pm, pp);
        const double f = e == 0.0 ? 0.0 : pm / e;
        se[0] = f * d[0];
        se[1] = f * d[1];
        se[2] = -2.0 * sqrt(pp);
    }

    // STICKY: update_z of the persistent shifted momentum to cos(z - A) = cz and
    // its rescue (PiFlux.cpp:253-257); returns its cos z
    __device__ __forceinline__ double shift_frame(double cz) const {
        if (ref_is_zero(eq_cx) && ref_is_zero(cy) && cz < 0.0) {
            eq_cx = fabs(cz) * kRescue;
            cz *= (1.0 - 0.5 * kRescue * kRescue);
        }
        last_cz = cz;
        return cz;
    }

    // STICKY: the shifted momentum's first update at A(t_begin), before the
    // integration (PiFlux.cpp:285)
    __device__ __forceinline__ void begin_run(double A0) {
        eq_cx = cx;
        shift_frame(cos(kz - A0));
    }

    // STICKY: compute_occupations of PiFlux.cpp:228-236 with the persistent
    // shifted momentum: the state in its band frame (diagonal_sigma,
    // PiFlux.cpp:1037-1049) d, then (lower, upper) = sqrt((|d| -+ d_z) / 2)
    __device__ __forceinline__ void occupations(const double s[3], double o[2]) const {
        const double mx = eq_cx, my = cy, mz = last_cz;
        const double disp = sqrt(mx * mx + my * my + mz * mz);
        const double z = mz + disp, nrm = 2.0 * disp * z;
        const double xz = 2.0 * mx * z, yz = 2.0 * my * z, xy = 2.0 * mx * my;
        const double d0 = (s[0] * (z * z - mx * mx + my * my) - xy * s[1] - xz * s[2]) / nrm;
        const double d1 = (s[1] * (z * z + mx * mx - my * my) - xy * s[0] - yz * s[2]) / nrm;
        const double d2 = (s[2] * (z * z - mx * mx - my * my) + xz * s[0] + yz * s[1]) / nrm;
        const double n = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        o[0] = sqrt(0.5 * (-d2 + n));
        o[1] = sqrt(0.5 * (d2 + n));
    }

    // ds/dt = 2 d(k - A) x s - relaxation
    __device__ __forceinline__ void rhs_A(const double s[3], double A, double out[3]) const {
        if constexpr (model_has_phase(M)) {
            double sn = 0.0, cs;
            if constexpr (M == MODEL_HONEYCOMB) sincos(0.5 * (kx - A), &sn, &cs);
            else cs = cos((M == MODEL_PIFLUX3D ? kz : kx) - A);
            rhs_sc(s, sn, cs, out);
            return;
        }
        if constexpr (M != MODEL_DIRAC3D && R != RELAX_EQUILIBRIUM) {
            // w = 2 d(k - A) with the factor 2 folded into the model's constant:
            // (2c) x equals 2 (c x) exactly, and the T2 projection of s on w equals
            // the one on d exactly (every product and sum scales by a power of
            // two), so this is the general path below with up to three
            // multiplications fewer per evaluation
            const double tv = 2.0 * J.v;
            const double wx = tv * (kx - A * J.ct), wy = tv * (ky - A * J.st), wz = 2.0 * J.mass;
            rhs_w(s, wx, wy, wz, out);
            return;
        }
        double d[3];
        d_of(A, d);
        const double wx = 2.0 * d[0], wy = 2.0 * d[1], wz = 2.0 * d[2];
        out[0] = wy * s[2] - wz * s[1];
        out[1] = wz * s[0] - wx * s[2];
        out[2] = wx * s[1] - wy * s[0];
        if constexpr (R == RELAX_TO_INITIAL) {
            const double r = J.relax_rate;
            out[0] -= (s[0] - s0[0]) * r;
            out[1] -= (s[1] - s0[1]) * r;
            out[2] -= (s[2] - s0[2]) * r;
        } else if constexpr (R == RELAX_T2) {
            const double r = J.relax_rate;
            const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            // 1/dd: hardware reciprocal refined by two Newton steps (within an
            // ulp of the quotient, fewer instructions than an IEEE division)
            double y = __builtin_amdgcn_rcp(dd);
            y = fma(fma(-dd, y, 1.0), y, y);
            y = fma(fma(-dd, y, 1.0), y, y);
            const double proj = dd > 0.0 ? (d[0] * s[0] + d[1] * s[1] + d[2] * s[2]) * y : 0.0;
            out[0] -= (s[0] - proj * d[0]) * r;
            out[1] -= (s[1] - proj * d[1]) * r;
            out[2] -= (s[2] - proj * d[2]) * r;
        }
    }

    // phase models at a stage whose field phase is (sin f A, cos f A)
    __device__ __forceinline__ void rhs_phase(const double s[3], double sa, double ca, double out[3]) const {
        double sn, cs;
        shifted_phase(sa, ca, sn, cs);
        rhs_sc(s, sn, cs, out);
    }

    // w x s - relaxation for w = 2 d; relaxation toward the initial state or T2
    __device__ __forceinline__ void rhs_w(const double s[3], double wx, double wy, double wz, double out[3]) const {
        out[0] = wy * s[2] - wz * s[1];
        out[1] = wz * s[0] - wx * s[2];
        out[2] = wx * s[1] - wy * s[0];
        if constexpr (R == RELAX_TO_INITIAL) {
            const double r = J.relax_rate;
            out[0] -= (s[0] - s0[0]) * r;
            out[1] -= (s[1] - s0[1]) * r;
            out[2] -= (s[2] - s0[2]) * r;
        } else if constexpr (R == RELAX_T2) {
            const double r = J.relax_rate;
            const double ww = wx * wx + wy * wy + wz * wz;
            double y = __builtin_amdgcn_rcp(ww);
            y = fma(fma(-ww, y, 1.0), y, y);
            y = fma(fma(-ww, y, 1.0), y, y);
            const double proj = ww > 0.0 ? (wx * s[0] + wy * s[1] + wz * s[2]) * y : 0.0;
            out[0] -= (s[0] - proj * wx) * r;
            out[1] -= (s[1] - proj * wy) * r;
            out[2] -= (s[2] - proj * wz) * r;
        }
    }

    // rhs_w for an in-plane field w = (wx, wy, 0) (the 2D lattice models): the
    // products with wz = 0, which IEEE arithmetic does not let the compiler
    // drop, left out (two operations per evaluation fewer)
    __device__ __forceinline__ void rhs_w0(const double s[3], double wx, double wy, double out[3]) const {
        out[0] = wy * s[2];
        out[1] = -(wx * s[2]);
        out[2] = wx * s[1] - wy * s[0];
        if constexpr (R == RELAX_TO_INITIAL) {
            const double r = J.relax_rate;
            out[0] -= (s[0] - s0[0]) * r;
            out[1] -= (s[1] - s0[1]) * r;
            out[2] -= (s[2] - s0[2]) * r;
        } else if constexpr (R == RELAX_T2) {
            const double r = J.relax_rate;
            const double ww = wx * wx + wy * wy;
            double y = __builtin_amdgcn_rcp(ww);
            y = fma(fma(-ww, y, 1.0), y, y);
            y = fma(fma(-ww, y, 1.0), y, y);
            const double proj = ww > 0.0 ? (wx * s[0] + wy * s[1]) * y : 0.0;
            out[0] -= (s[0] - proj * wx) * r;
            out[1] -= (s[1] - proj * wy) * r;
            out[2] -= s[2] * r;
        }
    }

    // phase models: sn, cs = sin, cos of f (k - A) (pi-flux: only cs is read)
    __device__ __forceinline__ void rhs_sc(const double s[3], double sn, double cs, double out[3]) const {
        if constexpr (M == MODEL_PIFLUX3D) {
            // 4t (m x s) (PiFlux.cpp:89-94), cz = cos(kz - A)
            const double cz = cs, pf = 4.0 * J.hop;
            out[0] = (cy * s[2] - cz * s[1]) * pf;
            out[1] = (cz * s[0] - cx * s[2]) * pf;
            out[2] = (cx * s[1] - cy * s[0]) * pf;
            if constexpr (R == RELAX_EQUILIBRIUM) relax_equilibrium(cz, s, out);
        } else if constexpr (M == MODEL_PIFLUX2D) {
            rhs_w0(s, (4.0 * J.hop) * cs, 2.0 * cy, out);
        } else if constexpr (R != RELAX_EQUILIBRIUM) {
            // honeycomb, w = 2 d = th (Re g, -Im g), th = 2t, g = e^{i qx} + 2 cy e^{-i qx/2}:
            // with a = th cos(qx/2), b = th sin(qx/2),
            //   th Re g = a (cos(qx/2) + 2 cy) - b sin(qx/2),  -th Im g = b (2 cy - 2 cos(qx/2))
            const double th = 2.0 * J.hop, a = th * cs, b = th * sn, c2 = 2.0 * cy;
            rhs_w0(s, fma(a, cs + c2, -(b * sn)), b * fma(-2.0, cs, c2), out);
        } else {
            // honeycomb: -(s - s_eq(k - A))/tau (Honeycomb.cpp:180-191)
            double d[3];
            honeycomb_d(sn, cs, d);
            const double wx = 2.0 * d[0], wy = 2.0 * d[1], wz = 2.0 * d[2];
            out[0] = wy * s[2] - wz * s[1];
            out[1] = wz * s[0] - wx * s[2];
            out[2] = wx * s[1] - wy * s[0];
            const double r = J.relax_rate;
            double se[3];
            honeycomb_target(d, se);
            out[0] -= r * (s[0] - se[0]);
            out[1] -= r * (s[1] - se[1]);
            out[2] -= r * (s[2] - se[2]);
        }
    }

    // 3D pi-flux: the instantaneous dispersion |m(k - A)| (PiFlux.cpp:400-402, 1001)
    __device__ __forceinline__ double dispersion_at(double A) const {
        const double c = cos(kz - A);
        return sqrt(cx * cx + cy * cy + c * c);
    }
    __device__ __forceinline__ double dispersion_phase(double sa, double ca) const {
        double sn, c;
        shifted_phase(sa, ca, sn, c);
        return sqrt(cx * cx + cy * cy + c * c);
    }

    // phase models: the current vertex at a sample whose field phase is (sin f A, cos f A)
    __device__ __forceinline__ double vertex_phase(double sa, double ca, const double s[3]) const {
        double sn, cs;
        shifted_phase(sa, ca, sn, cs);
        if constexpr (M == MODEL_PIFLUX3D) {
            return sn * s[2];  // sin(kz - A) s_z
        } else if constexpr (M == MODEL_PIFLUX2D) {
            return sn * s[0];  // sin(kx - A) s_x
        } else {
            const double cq = (cs - sn) * (cs + sn), sq = 2.0 * sn * cs;
            return (cq - cs * cy) * s[0] + (sq + sn * cy) * s[1];
        }
    }

    // honeycomb j_perp vertex (the y component at theta = 0) divided by
    // perp_factor(): with G_y = sum_delta delta_y e^{i q.delta}
    // = i sqrt3 sin(sqrt3/2 ky) e^{-i qx/2}, Re G_y s_x + Im G_y s_y
    // = sqrt3 sin(sqrt3/2 ky) (sin(qx/2) s_x + cos(qx/2) s_y)
    __device__ __forceinline__ double vertex_perp_phase(double sa, double ca, const double s[3]) const {
        double sn, cs;
        shifted_phase(sa, ca, sn, cs);
        return sn * s[0] + cs * s[1];
    }
    __device__ __forceinline__ double perp_factor() const { return kSqrt3 * sin(kHalfSqrt3 * ky); }

    // current vertex g(s, A) along the polarisation
    __device__ __forceinline__ double vertex(double A, const double s[3]) const {
        if constexpr (M == MODEL_PIFLUX3D) {
            return sin(kz - A) * s[2];  // PiFlux.cpp:831-838
        } else if constexpr (M == MODEL_DIRAC3D) {
            return s[2];  // band inversion; the weight carries k_z kappa/|k| (DiracSystem.cpp:279)
        } else if constexpr (M == MODEL_DIRAC2D) {
            return -J.v * (J.ct * s[0] + J.st * s[1]);
        } else if constexpr (M == MODEL_PIFLUX2D) {
            return sin(kx - A) * s[0];
        } else {
            // Gx = sum_delta delta_x e^{i q.delta} = e^{i qx} - e^{-i qx/2} cos(sqrt3/2 qy)
            // (shifted_grad_gamma, Honeycomb.cpp:52-61), contracted as Honeycomb.cpp:273-279
            double sh, ch;
            sincos(0.5 * (kx - A), &sh, &ch);
            const double cq = (ch - sh) * (ch + sh), sq = 2.0 * sh * ch;
            return (cq - ch * cy) * s[0] + (sq + sh * cy) * s[1];
        }
    }
};

// ---------------------------------------------------------------------------
// Fehlberg 7(8) (E. Fehlberg, NASA TR R-287, 1968).  Stages are fully
// unrolled so every coefficient is an immediate and zero entries vanish.
// ---------------------------------------------------------------------------
struct F78 {
    static constexpr double c[13] = {0.0,       2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 1.0 / 2.0, 5.0 / 6.0,
                                     1.0 / 6.0, 2.0 / 3.0,  1.0 / 3.0, 1.0,       0.0,        1.0};
    static constexpr double a[13][12] = {
        {0},
        {2.0 / 27.0},
        {1.0 / 36.0, 1.0 / 12.0},
        {1.0 / 24.0, 0.0, 1.0 / 8.0},
        {5.0 / 12.0, 0.0, -25.0 / 16.0, 25.0 / 16.0},
        {1.0 / 20.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 5.0},
        {-25.0 / 108.0, 0.0, 0.0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0},
        {31.0 / 300.0, 0.0, 0.0, 0.0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0},
        {2.0, 0.0, 0.0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0},
        {-91.0 / 108.0, 0.0, 0.0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0, 17.0 / 6.0,
         -1.0 / 12.0},
        {2383.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0, 45.0 / 82.0,
         45.0 / 164.0, 18.0 / 41.0},
        {3.0 / 205.0, 0.0, 0.0, 0.0, 0.0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0, 0.0},
        {-1777.0 / 4100.0, 0.0, 0.0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0, 51.0 / 82.0,
         33.0 / 164.0, 12.0 / 41.0, 0.0, 1.0},
    };
    // 8th-order (propagated) weights; the embedded 7th-order solution differs
    // by 41/840 (k1 + k11 - k12 - k13).
    static constexpr double b8[13] = {0.0,         0.0,         0.0, 0.0,          0.0,         34.0 / 105.0, 9.0 / 35.0,
                                      9.0 / 35.0,  9.0 / 280.0, 9.0 / 280.0, 0.0, 41.0 / 840.0, 41.0 / 840.0};
};

struct StepResult {
    double err;  // max_i |err_i| / (atol + rtol (|x_i| + h |dxdt_i|))
};

// One embedded RKF78 attempt from (s, t) with step h; k1 = f(s, t) given.
// Largest step growth of the controller: 0.9 * (5^-8)^(-1/8), evaluated with
// the same operations as the accept branch of rkf78_interval so that replayed
// nominal steps are bitwise equal to the integrator's.
__device__ __forceinline__ double max_growth() {
    const double e = 1.0 / 390625.0;
    return 0.9 / sqrt(sqrt(sqrt(e)));
}

// The accepted step's growth: for err < 0.5, h 0.9 err^(-1/8), at most
// max_growth().
__device__ __forceinline__ void grow_step(double& h, double err) {
    if (err < 0.5) {
        if (err > (1.0 / 390625.0)) h *= 0.9 / sqrt(sqrt(sqrt(err)));
        else h *= max_growth();
    }
}

// The pulse does not depend on the state, so the field at the stage times of
// a step is produced first: from the nominal-step table row `cand` when its
// (t, h) equal the step's, else from the per-step pulse cache.  All the row's
// words are loaded at once so the match test and the values arrive together.
// Stage values: A at the 13 stage times, or for the phase models
// (sin f A, cos f A) at the 10 distinct stage times (model_has_phase).
template <bool PHASE>
struct StageVals {
    double A[13];
};
template <>
struct StageVals<true> {
    double S[kPhaseStages], C[kPhaseStages];
};

// The stage values of a table row (words 2..W-1).
template <class F>
__device__ __forceinline__ void stage_values_from_row(const double* cv, StageVals<F::kPhase>& sv) {
    if constexpr (F::kPhase) {
#pragma unroll
        for (int d = 0; d < kPhaseStages; ++d) sv.S[d] = cv[2 + 2 * d], sv.C[d] = cv[3 + 2 * d];
    } else {
#pragma unroll
        for (int j = 0; j < 13; ++j) sv.A[j] = cv[2 + j];
    }
}

// The stage values of the step (t, h) from the per-step pulse cache.  The
// pulse parameters are read afresh at every step (fresh_job).
//
// Phase models: sin and cos of f A at the first stage time (c = 0) in full;
// the other stage times follow by angle addition, sin(x0 + e) = sin x0 cos e
// + cos x0 sin e with e = f (A_d - A_0) and small_sincos(e).  Where an e of an
// active lane is outside small_sincos's range, the active lanes evaluate every
// stage's sincos in full instead (a branch uniform over them).  Both forms
// agree to a few ulp.
template <class F>
__device__ __forceinline__ void stage_values_from_cache(F& f, double t, double h, StageVals<F::kPhase>& sv) {
    JobRef J = fresh_job(f.J);
    f.pc.prepare(J, t);
    if constexpr (F::kPhase) {
        constexpr double fs = model_phase_scale(F::kModel);
        static_assert(F78::c[phase_stage_of(0)] == 0.0, "the first distinct stage time is the step's start");
        const double x0 = fs * f.pc.at(J, 0.0);
        sincos(x0, &sv.S[0], &sv.C[0]);
        bool small = true;
#pragma unroll
        for (int d = 1; d < kPhaseStages; ++d) {
            const double e = fs * f.pc.at(J, F78::c[phase_stage_of(d)] * h) - x0;
            small = small && fabs(e) <= kSmallArg;
            double se, ce;
            small_sincos(e, se, ce);
            sv.S[d] = sv.S[0] * ce + sv.C[0] * se;
            sv.C[d] = sv.C[0] * ce - sv.S[0] * se;
        }
        // marked unlikely: placed off the straight-line path, the full form
        // leaves the surrounding loop's code as it is without it
        if (__builtin_expect(!__all(small), 0)) {
#pragma unroll
            for (int d = 1; d < kPhaseStages; ++d)
                sincos(fs * f.pc.at(J, F78::c[phase_stage_of(d)] * h), &sv.S[d], &sv.C[d]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 13; ++i) {
            if (i == 7) sv.A[i] = sv.A[3];
            else if (i == 11) sv.A[i] = sv.A[0];
            else if (i == 12) sv.A[i] = sv.A[10];
            else sv.A[i] = f.pc.at(J, F78::c[i] * h);
        }
    }
}

template <class F>
__device__ __forceinline__ void stage_pulses(F& f, double t, double h, const double* cand,
                                             StageVals<F::kPhase>& sv) {
    constexpr int W = nom_stride(F::kModel, F::kRelax);
    bool hit = false;
    if (cand) {
        double cv[W];
#pragma unroll
        for (int j = 0; j < W; ++j) cv[j] = cand[j];
        hit = (cv[0] == t) && (cv[1] == h);
        stage_values_from_row<F>(cv, sv);
    }
    if (!hit) stage_values_from_cache(f, t, h, sv);
}

// right-hand side at stage i
template <class F>
__device__ __forceinline__ void stage_rhs(const F& f, const StageVals<F::kPhase>& sv, int i, const double x[3],
                                          double out[3]) {
    if constexpr (F::kPhase) f.rhs_phase(x, sv.S[phase_stage(i)], sv.C[phase_stage(i)], out);
    else f.rhs_A(x, sv.A[i], out);
}

// Like odeint's try_step for a non-FSAL stepper, each attempt evaluates the
// derivative at its start, so nothing but the state crosses attempts.
// Below this ratio of |err_i| to its scale every component's error ratio rounds
// to at most the controller's growth threshold 5^-8 (the factor 1 - 2^-50 covers
// the rounding of the product and of the quotient).
constexpr double kTinyErrRatio = (1.0 / 390625.0) * (1.0 - 0x1p-50);

// tiny_mode TINY_ZERO: when every component has scale >= 2^-1000 (so
// scale * ratio is a normal number) and |err_i| < kTinyErrRatio * scale, return
// 0 without the divisions; the exact ratio is then at most 5^-8 and the
// controller takes the same branch (accept, largest growth) whatever its value.
// Only step logs read the value below 5^-8 (the network's feature clamps the
// ratio at 5^-8).  NaN, infinite and zero-scale cases fail the test and take
// the exact path.  TINY_EXACT always computes the exact ratio.
enum : int { TINY_EXACT = 0, TINY_ZERO = 1 };

template <class F>
__device__ __forceinline__ double rkf78_attempt(F& f, const double s[3], double h, const StageVals<F::kPhase>& sv,
                                                double snew[3], double rtol, double atol, int tiny_mode = TINY_EXACT) {
    double k[13][3];
    f.begin_step(s);
    stage_rhs(f, sv, 0, s, k[0]);
    const double k1[3] = {k[0][0], k[0][1], k[0][2]};
#pragma unroll
    for (int i = 1; i < 13; ++i) {
        double x[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < i; ++j)
                if (F78::a[i][j] != 0.0) acc = fma(F78::a[i][j], k[j][q], acc);
            x[q] = fma(h, acc, s[q]);
        }
        stage_rhs(f, sv, i, x, k[i]);
    }
    double e[3], sc[3];
    bool tiny = tiny_mode != TINY_EXACT;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 13; ++j)
            if (F78::b8[j] != 0.0) acc = fma(F78::b8[j], k[j][q], acc);
        snew[q] = fma(h, acc, s[q]);
        e[q] = h * (41.0 / 840.0) * ((k[0][q] + k[10][q]) - (k[11][q] + k[12][q]));
        sc[q] = atol + rtol * (fabs(s[q]) + fabs(h) * fabs(k1[q]));
        tiny = tiny && sc[q] >= 0x1p-1000 && fabs(e[q]) < sc[q] * kTinyErrRatio;
    }
    // a wave-uniform branch: with a per-lane one the compiler computes the
    // divisions for every lane and selects
    if (__all(tiny)) return 0.0;
    double err = 0.0;
    bool nan = false;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double r = fabs(e[q]) / sc[q];
        nan |= (r != r);
        err = r > err ? r : err;
    }
    return tiny ? 0.0 : nan ? __builtin_nan("") : err;
}

// Options and results of one sampling interval beyond the odeint semantics.
struct IntervalCtl {
    int fallback_after;  // rejections of the first step after which it restarts at h_fallback (0: never)
    double h_fallback;   // the plain controller's restart step Delta / n_sub
    int tiny_mode;       // rkf78_attempt's TINY_*: TINY_ZERO unless a step log reads the error ratios
};
struct IntervalOut {
    double h_acc;        // size of the first accepted step of the interval (0: none yet)
    double err_acc;      // its error ratio
    double last_err;     // error ratio of the last accepted step
    uint32_t attempts;   // attempts in this interval
    bool fell_back;      // the first step was handed back to the plain controller
};

// integrate_adaptive over [t0, t1] (odeint semantics: absolute-epsilon loop
// and clip tests, reject -> dt *= max(0.9 err^(-1/6), 0.2), accept and for
// err < 0.5 grow dt *= 0.9 max(err, 5^-8)^(-1/8)).  h is the starting step
// and on return the controller's proposal; unclipped the step that was in
// force before the final clip.  status: 0 ok, 1 controller gave up, 2 the
// state became non-finite.
//
// Forward-progress guard (not in odeint, whose loop would repeat such steps
// forever): the lane fails with status 1 when an accepted step leaves t
// unchanged (h below half an ulp of t) or when the interval takes more than
// sp.max_steps accepted steps.  Every other loop exit is odeint's, so every
// lane leaves this loop after at most max_steps accepted and max_fail + 1
// attempts per step.
//
// nom: this interval's row of the nominal-step table (nullptr: none).
template <class F>
__device__ __forceinline__ void rkf78_interval(F& f, double s[3], double t0, double t1, double& h,
                                               double& unclipped, const SolverDev& sp, uint32_t& n_att,
                                               uint32_t& n_rej, int& status, const double* nom,
                                               const IntervalCtl& ctl, IntervalOut& io) {
    double t = t0;
    bool fresh = true;
    int tries = 0;
    int step = -1;  // index of the current step inside the interval
    io.attempts = 0;
    io.fell_back = false;
    io.h_acc = 0.0;
    io.err_acc = 0.0;
    for (;;) {
        if (fresh) {
            if (!(t1 - t > kDblEps)) break;
            if (step + 1 >= sp.max_steps) {
                status = 1;
                break;
            }
            unclipped = h;
            if ((t + h) - t1 > kDblEps) h = t1 - t;
            tries = 0;
            fresh = false;
            ++step;
        }
        // candidate slot: the restart sequence's step, or the whole-interval
        // step for a carried or proposed first step
        constexpr int W = nom_stride(F::kModel, F::kRelax);
        const double* cand = nullptr;
        if (nom) {
            if (sp.proposal == PROPOSAL_RESTART || io.fell_back) {
                if (step < kNomRestartSlots) cand = nom + step * W;
            } else if (step == 0) {
                cand = nom + kNomRestartSlots * W;
            }
        }
        StageVals<F::kPhase> sv;
        stage_pulses(f, t, h, cand, sv);
        double snew[3];
        const double err = rkf78_attempt(f, s, h, sv, snew, sp.rtol, sp.atol, ctl.tiny_mode);
        ++n_att;
        ++io.attempts;
        if (err > 1.0) {
            ++n_rej;
            double fac = 0.9 * pow(err, -1.0 / 6.0);
            fac = fac > 0.2 ? fac : 0.2;
            h *= fac;
            ++tries;
            if (ctl.fallback_after > 0 && step == 0 && !io.fell_back && tries >= ctl.fallback_after) {
                h = ctl.h_fallback;
                io.fell_back = true;
            }
            if (tries > sp.max_fail || !(h > 0.0)) {
                status = 1;
                break;
            }
        } else {
            const double t_next = t + h;
            if (!(t_next > t)) {
                status = 1;
                break;
            }
            t = t_next;
            s[0] = snew[0];
            s[1] = snew[1];
            s[2] = snew[2];
            io.last_err = err;
            if (io.h_acc == 0.0) io.h_acc = h, io.err_acc = err;
            grow_step(h, err);
            fresh = true;
            if (!(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]))) {
                status = 2;
                break;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Rotations s <- exp(W) s, W v = O x v: Rodrigues' formula with the half-angle
// form of 1 - cos and a small-angle series (cf. the reference's closed form,
// sources/HHG/Systems/DiracDetail/MagnusMatrix.cpp:64-74, which divides by
// theta^2 directly).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rotate_about(const double O[3], double s[3]) {
    const double th2 = O[0] * O[0] + O[1] * O[1] + O[2] * O[2];
    double sa, sb;  // sin(th)/th, (1 - cos th)/th^2
    if (th2 < 1e-2) {
        sa = 1.0 + th2 * (-1.0 / 6.0 + th2 * (1.0 / 120.0 + th2 * (-1.0 / 5040.0 + th2 * (1.0 / 362880.0 +
                                                                                            th2 * (-1.0 / 39916800.0)))));
        sb = 0.5 + th2 * (-1.0 / 24.0 + th2 * (1.0 / 720.0 + th2 * (-1.0 / 40320.0 + th2 * (1.0 / 3628800.0 +
                                                                                             th2 * (-1.0 / 479001600.0)))));
    } else {
        const double th = sqrt(th2);
        double sn, cs;
        sincos(0.5 * th, &sn, &cs);
        sa = 2.0 * sn * cs / th;
        sb = 2.0 * sn * sn / th2;
    }
    const double x0 = O[1] * s[2] - O[2] * s[1];
    const double x1 = O[2] * s[0] - O[0] * s[2];
    const double x2 = O[0] * s[1] - O[1] * s[0];
    const double y0 = O[1] * x2 - O[2] * x1;
    const double y1 = O[2] * x0 - O[0] * x2;
    const double y2 = O[0] * x1 - O[1] * x0;
    s[0] += sa * x0 + sb * y0;
    s[1] += sa * x1 + sb * y1;
    s[2] += sa * x2 + sb * y2;
}

// Magnus 4 (two Gauss nodes): Omega = h/2 (w1 + w2) + sqrt3/12 h^2 (w2 x w1), w = 2 d.
template <class F>
__device__ __forceinline__ void magnus4_step(const F& f, double s[3], double A1, double A2, double h) {
    double w1[3], w2[3];
    f.d_of(A1, w1);
    f.d_of(A2, w2);
    const double a = h, b = (kSqrt3 / 12.0) * 4.0 * h * h;  // omega = 2 d -> w2 x w1 scales by 4
    double O[3];
    O[0] = a * (w1[0] + w2[0]) + b * (w2[1] * w1[2] - w2[2] * w1[1]);
    O[1] = a * (w1[1] + w2[1]) + b * (w2[2] * w1[0] - w2[0] * w1[2]);
    O[2] = a * (w1[2] + w2[2]) + b * (w2[0] * w1[1] - w2[1] * w1[0]);
    rotate_about(O, s);
}

// ---------------------------------------------------------------------------
// The reference's Magnus step of the 3D Dirac cone, one per sampling interval
// (DiracSystem::time_evolution_magnus, DiracSystem.cpp:176-211).  Per lane:
// K = 2|k| h, Kp = 2 kappa h, Kz = 2 k_z h (Magnus(2|k|, 2 kappa, 2 k_z, h),
// MagnusMatrix.cpp:12-23); per interval the four pulse moments divided by |k|
// (al, be, ga, de).  Omega = [[0, a, b], [-a, 0, c], [-b, -c, 0]] with the
// entries of MagnusMatrix.cpp:35-62, grouped here by q = Kz^2 + Kp^2; then
// s <- exp(Omega) s, a rotation about (-c, b, -a).
// ---------------------------------------------------------------------------
struct MomentLane {
    double K, K2, K3, K4, Kp, Kp2, Kz, Kz2, q, ik;
};

__device__ __forceinline__ MomentLane moment_lane(double kz, double kappa, double h) {
    MomentLane L;
    const double kk = sqrt(kz * kz + kappa * kappa);
    L.ik = 1.0 / kk;
    L.K = 2.0 * kk * h;
    L.K2 = L.K * L.K;
    L.K3 = L.K2 * L.K;
    L.K4 = L.K2 * L.K2;
    L.Kp = 2.0 * kappa * h;
    L.Kp2 = L.Kp * L.Kp;
    L.Kz = 2.0 * kz * h;
    L.Kz2 = L.Kz * L.Kz;
    L.q = L.Kz2 + L.Kp2;
    return L;
}

__device__ __forceinline__ void magnus_moments_step(const MomentLane& L, const double mom[4], double s[3]) {
    const double al = mom[0] * L.ik, be = mom[1] * L.ik, ga = mom[2] * L.ik, de = mom[3] * L.ik;
    const double K = L.K, K2 = L.K2, K3 = L.K3, K4 = L.K4, Kp = L.Kp, Kp2 = L.Kp2, Kz = L.Kz, Kz2 = L.Kz2, q = L.q;
    const double al2 = al * al, al3 = al2 * al, al4 = al2 * al2, be2 = be * be, be3 = be2 * be, ga2 = ga * ga;
    const double a = al * Kz - K +
                     Kp2 * ((al2 * be2 - al3 * ga) * K * q * (1.0 / 2520.0) + al2 * ga * K2 * Kz * (1.0 / 1260.0) -
                            al * be2 * K2 * Kz * (1.0 / 1080.0) - al * ga * K3 * (1.0 / 2520.0) +
                            be2 * K3 * (1.0 / 1890.0) + be2 * K * (1.0 / 60.0) - be * de * K * (1.0 / 420.0) +
                            ga2 * K * (1.0 / 420.0));
    const double b = K * Kp *
                     (-al4 * be * q * q + 4.0 * al3 * be * K * Kz * q - 6.0 * al2 * be * K2 * Kz2 -
                      2.0 * al2 * be * K2 * Kp2 + (-42.0 * al2 * be + 18.0 * al2 * de - 36.0 * al * be * ga + 18.0 * be3) * q +
                      (4.0 * al * be * K3 + 84.0 * al * be * K - 36.0 * al * de * K + 36.0 * be * ga * K) * Kz -
                      be * K4 - 42.0 * be * K2 - 2520.0 * be + 18.0 * de * K2) *
                     (1.0 / 15120.0);
    const double c = Kp *
                     (3.0 * (al2 * be2 - al3 * ga) * K * Kz * q + 9.0 * al2 * ga * K2 * Kz2 + 3.0 * al2 * ga * K2 * Kp2 -
                      6.0 * al * be2 * K2 * Kz2 + al * be2 * K2 * Kp2 - 9.0 * al * ga * K3 * Kz - 7560.0 * al +
                      (3.0 * be2 * K3 + 126.0 * be2 * K - 18.0 * be * de * K + 18.0 * ga2 * K) * Kz + 3.0 * ga * K4) *
                     (1.0 / 7560.0);
    const double O[3] = {-c, b, -a};
    rotate_about(O, s);
}

// ---------------------------------------------------------------------------
// Wavefront reduction and exact cross-wave accumulation.
// ---------------------------------------------------------------------------
// Butterfly over 64 lanes; IEEE addition is commutative, so every lane ends
// with the bitwise-identical total.
// v plus the value DPP control CTRL moves into each lane (rows outside ROWS
// add 0); the two 32-bit halves move with the same control
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWS, 0xf, false);
    return v + __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a fully active wave, returned wave-uniform: a DPP
// tree (pairs quad_perm [1,0,3,2], quads [2,3,0,1], half rows row_half_mirror,
// rows row_mirror, then row_bcast:15 into rows 1 and 3 and row_bcast:31 into
// rows 2 and 3) leaves the total in lane 63, which v_readlane returns.  All
// moves are VALU operations: no LDS round trips on the per-sample path.  The
// order of the additions is fixed, so the sum is deterministic.  The rows a
// broadcast leaves out keep whatever the moved register held (no zeroing
// move): they add an undefined value to rows 0 and 2, which lane 63's sum
// never reads (row_bcast:15 reads lane 15 and row_bcast:31 lane 31 before the
// addition of their step).
__device__ __forceinline__ double wave_sum(double v) {
    v = dpp_add<0xb1, 0xf>(v);
    v = dpp_add<0x4e, 0xf>(v);
    v = dpp_add<0x141, 0xf>(v);
    v = dpp_add<0x140, 0xf>(v);
    v = dpp_add<0x142, 0xa>(v);
    v = dpp_add<0x143, 0xc>(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Signed 128-bit fixed point with 80 fractional bits: x -> round-toward-minus
// of x 2^80 split as hi 2^64 + lo.  Integer addition is associative, so the
// accumulated j is bitwise independent of the order in which waves finish.
constexpr double kFixScaleHi = 65536.0;                 // 2^16
constexpr double kFixScaleLo = 18446744073709551616.0;  // 2^64
constexpr double kFixLimit = 4611686018427387904.0;     // 2^62 (|x 2^16| bound)

__device__ __forceinline__ bool fix_split(double x, unsigned long long& lo, unsigned long long& hi) {
    const double y = x * kFixScaleHi;
    if (!(fabs(y) < kFixLimit)) return false;
    const double hi_d = floor(y);
    const double frac = y - hi_d;  // exact, in [0, 1)
    lo = (unsigned long long)(frac * kFixScaleLo);
    hi = (unsigned long long)(long long)hi_d;
    return true;
}

// acc (2 words) += (lo, hi) as one 128-bit two's-complement addition; the
// carry comes from the low word's old value.  Addition modulo 2^128 is
// associative, so partial sums may be added in any grouping and order.
__device__ __forceinline__ void fix_add(unsigned long long* acc, unsigned long long lo, unsigned long long hi) {
    const unsigned long long old = atomicAdd(acc, lo);
    const unsigned long long carry = (old + lo < old) ? 1ull : 0ull;
    atomicAdd(acc + 1, hi + carry);
}

// fix_add on the block's LDS copy (workgroup-scope atomics)
__device__ __forceinline__ void fix_add_lds(unsigned long long* acc, unsigned long long lo, unsigned long long hi) {
    const unsigned long long old = __hip_atomic_fetch_add(acc, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const unsigned long long carry = (old + lo < old) ? 1ull : 0ull;
    __hip_atomic_fetch_add(acc + 1, hi + carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool fix_accumulate(unsigned long long* acc, double x) {
    unsigned long long lo, hi;
    if (!fix_split(x, lo, hi)) return false;
    fix_add(acc, lo, hi);
    return true;
}

__device__ __forceinline__ double fix_to_double(const unsigned long long* acc) {
    const long long hi = (long long)acc[1];
    const unsigned long long lo = acc[0];
    return (double)hi * (1.0 / kFixScaleHi) + (double)lo * (1.0 / (kFixScaleHi * kFixScaleLo));
}

}  // namespace hhg
