// ftwa_internal.hpp -- helpers shared by the C ABI implementation and the JSON front end.
#pragma once

#include <string>

namespace fpsamd {

// Time-dependent pulse amplitude A(t) (multiplies the polarisation vector).
struct PulseSpec {
    int type;  // FPSFTWA_PULSE_*
    double A0, t0, td, w0, phi;
    const double* knots;
    int n_knots;
    double t_max;
};
double pulse_amplitude(const PulseSpec& p, double t);

void set_last_error(const std::string& msg);

}  // namespace fpsamd
