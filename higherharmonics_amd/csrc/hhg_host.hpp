// hhg_host.hpp — host-side helpers shared by the C ABI translation units.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hhg_amd.h"

namespace hhg {

struct MlpDev;
const MlpDev* mlp_device_record(const hhg_mlp* m);
int mlp_device(const hhg_mlp* m);

hhg_status fail(hhg_status st, const std::string& msg);
void clear_error();
const char* last_error();

void gauss_legendre(int n, double* x, double* w);
int64_t kgrid_size(const hhg_kgrid_desc& g);
hhg_status kgrid_build(const hhg_kgrid_desc& g, double* kx, double* ky, double* kz, double* w);
double pulse_eval(const hhg_pulse_desc& p, double t);  // analytic kinds
hhg_status spline_coefficients(const double* y, int64_t n, double h, double d0, double d1, double* coef);

// A spline pulse's coefficient table and its evaluation (HHG_PULSE_SPLINE).
struct SplinePulse {
    std::vector<double> coef;  // 4 per interval
    int64_t n = 0;             // intervals
    double t0 = 0.0, inv_h = 0.0, t_last = 0.0;
    hhg_status status = HHG_SUCCESS;
    explicit SplinePulse(const hhg_pulse_desc& p);
    double operator()(double t) const;
};
hhg_status validate_job(const hhg_job_desc& jd, const hhg_solver_desc& sv, int n_kgrids, int idx);
hhg_status trapezoidal_spectrum(const double* j, const hhg_time_desc& tc, int zero_padding, double* freq, double* re,
                                double* im);

}  // namespace hhg
