// Error norms and the tangency diagnostic on device fields: LPM's ErrNorms
// (src/lpm_error.hpp:81-131, src/lpm_error_impl.hpp:61-182) and
// SphereTangentFunctor (src/lpm_sphere_functions.hpp:62-78), over the C ABI
// (lpm_err_norms, lpm_sphere_tangency).
#pragma once

#include <cstdio>
#include <string>

#include "lpm.hpp"

namespace Lpm {

// Relative l1, l2 and linf errors of an approximation against exact
// values, weighted by `wt` (l1, l2), as the reference defines them.
struct ErrNorms {
  Real l1 = 0;
  Real l2 = 0;
  Real linf = 0;

  ErrNorms(Real l_1, Real l_2, Real l_i) : l1(l_1), l2(l_2), linf(l_i) {}

  // err = appx - exact (zero where mask is set; err may be nullptr), then
  // the norms.  Fields are device SoA arrays of n elements with ncomp
  // components at stride n (scalars: ncomp = 1); wt and mask may be
  // nullptr.  Blocks on the stream.
  ErrNorms(const Context& ctx, Real* err, const Real* appx, const Real* exact,
           const Real* wt, const uint8_t* mask, int64_t n, int ncomp,
           hipStream_t stream = nullptr) {
    double out[3];
    check(lpm_err_norms(ctx.get(), appx, 1, n, exact, 1, n, wt, mask, n,
                        ncomp, err, 1, n, out, stream),
          ctx.get());
    l1 = out[0];
    l2 = out[1];
    linf = out[2];
  }

  // Vector fields (three components at the field's component stride) over
  // their first n elements; err may be nullptr.
  ErrNorms(const Context& ctx, VectorField* err, const VectorField& appx,
           const VectorField& exact, const Real* wt, const uint8_t* mask,
           int64_t n, hipStream_t stream = nullptr) {
    double out[3];
    check(lpm_err_norms(ctx.get(), appx.data(), appx.es(), appx.cs(),
                        exact.data(), exact.es(), exact.cs(), wt, mask, n, 3,
                        err ? err->data() : nullptr, err ? err->es() : 1,
                        err ? err->cs() : 1, out, stream),
          ctx.get());
    l1 = out[0];
    l2 = out[1];
    linf = out[2];
  }

  std::string info_string(const std::string& label = "",
                          int tab_level = 0) const {
    std::string tabs(static_cast<size_t>(tab_level), '\t');
    char buf[160];
    std::snprintf(buf, sizeof buf, "%sErrNorms %s: l1 = %.6e, l2 = %.6e, linf = %.6e\n",
                  tabs.c_str(), label.c_str(), l1, l2, linf);
    return buf;
  }
};

// udotx[i] = u_i . x_i for the first n particles of two vector fields
// (udotx: device array of n values).
inline void sphere_tangency(const Context& ctx, Real* udotx,
                            const VectorField& x, const VectorField& u,
                            int64_t n, hipStream_t stream = nullptr) {
  check(lpm_sphere_tangency(ctx.get(), u.data(), u.es(), u.cs(), x.data(),
                            x.es(), x.cs(), n, udotx, stream),
        ctx.get());
}

}  // namespace Lpm
