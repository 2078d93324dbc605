// Host helpers shared by the renderer and the batch renderer.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/rayna_gpu.h"
#include "layout.hpp"

namespace rayna {

// Camera::calculate_viewport (scene/camera.rs:114-167) in double precision, rounded to float.
inline bool viewport(const rayna_camera &c, GViewport &vp) {
    const double focal = c.focus_dist;
    if (c.v_fov == 0.0 || focal == 0.0) return false;
    const double fl = std::sqrt((c.fwd[0] * c.fwd[0] + c.fwd[1] * c.fwd[1]) + c.fwd[2] * c.fwd[2]);
    const double rf = 1.0 / fl;
    if (!std::isfinite(rf) || !(rf > 0)) return false;
    const double w[3] = {-(c.fwd[0] * rf), -(c.fwd[1] * rf), -(c.fwd[2] * rf)};
    // u = cross(Y, w)
    const double cu[3] = {1.0 * w[2] - 0.0 * w[1], 0.0 * w[0] - 0.0 * w[2], 0.0 * w[1] - 1.0 * w[0]};
    const double ul = std::sqrt((cu[0] * cu[0] + cu[1] * cu[1]) + cu[2] * cu[2]);
    const double ru = 1.0 / ul;
    if (!std::isfinite(ru) || !(ru > 0)) return false;
    const double u[3] = {cu[0] * ru, cu[1] * ru, cu[2] * ru};
    const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
    const double defocus_radius = focal * std::tan(c.defocus_angle / 2.0);
    const double hh = std::tan(c.v_fov / 2.0);
    const double size = 2.0 * hh * focal;
    for (int i = 0; i < 3; ++i) {
        vp.pos[i] = (float)c.pos[i];
        vp.pixel_center[i] = (float)(c.pos[i] - w[i] * focal);
        vp.du[i] = (float)(u[i] * defocus_radius);
        vp.dv[i] = (float)(v[i] * defocus_radius);
        vp.vu[i] = (float)(u[i] * size);
        vp.vv[i] = (float)(-v[i] * size);
    }
    return true;
}

inline uint32_t isqrt(uint32_t n) {
    uint32_t s = (uint32_t)std::sqrt((double)n);
    while (s * s > n) --s;
    while ((s + 1) * (s + 1) <= n) ++s;
    return s;
}

}  // namespace rayna
