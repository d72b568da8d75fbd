// Per-node arithmetic shared by the node launch (node_kernels.hip) and the
// fused tile launch (tile_kernels.hip).  Contraction is off in every
// function: each statement rounds as the reference's C++ does, so given the
// same forces the update is bit-identical to the CPU path.
#pragma once

#include <math.h>

#include "device_types.h"

namespace ms {

// Damping term of one dof, evaluated as calc_non_deformation_forces.cpp:55-57.
__device__ __forceinline__ double damped(double f, double eta, double m, double v, double den) {
#pragma clang fp contract(off)
    double t = -eta * m;
    t = t * v;
    t = t / den;
    return f + t;
}

// The update of one dof (advance_dynamics.cpp:81-92).
__device__ __forceinline__ void update(const ShellDev& s, double m, double f, double& v, double& x) {
#pragma clang fp contract(off)
    if (!s.gd) {
        v = v + (s.dt / m) * f;
    } else {
        v = v + (s.dt / m) * f;
        v = (s.damp_den * f) / (s.eta * m);
    }
    x = x + s.dt * v;
}

__device__ __forceinline__ double norm3(double a, double b, double c) {
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}

// Adds contributions e0..e1-1 (stride 3 per entry, as (x, y, z) triples at
// base[idx(e)*3]) to the running sums, one at a time in the order given
// (four loads in flight), so a sum split over several calls rounds exactly
// as one call over the whole range.
template <typename CT, typename IDX>
__device__ __forceinline__ void ordered_sum3_acc(const CT* base, IDX idx, long long e0, long long e1, double& ax,
                                                 double& ay, double& az) {
#pragma clang fp contract(off)
    long long e = e0;
    for (; e + 4 <= e1; e += 4) {
        const CT* c0 = base + (size_t)idx(e) * 3;
        const CT* c1 = base + (size_t)idx(e + 1) * 3;
        const CT* c2 = base + (size_t)idx(e + 2) * 3;
        const CT* c3 = base + (size_t)idx(e + 3) * 3;
        const double x0 = c0[0], y0 = c0[1], z0 = c0[2], x1 = c1[0], y1 = c1[1], z1 = c1[2];
        const double x2 = c2[0], y2 = c2[1], z2 = c2[2], x3 = c3[0], y3 = c3[1], z3 = c3[2];
        ax = ax + x0;
        ay = ay + y0;
        az = az + z0;
        ax = ax + x1;
        ay = ay + y1;
        az = az + z1;
        ax = ax + x2;
        ay = ay + y2;
        az = az + z2;
        ax = ax + x3;
        ay = ay + y3;
        az = az + z3;
    }
    for (; e < e1; ++e) {
        const CT* c = base + (size_t)idx(e) * 3;
        ax = ax + (double)c[0];
        ay = ay + (double)c[1];
        az = az + (double)c[2];
    }
}

// The sum of contributions e0..e1-1 in the order given.
template <typename CT, typename IDX>
__device__ __forceinline__ void ordered_sum3(const CT* base, IDX idx, long long e0, long long e1, double& fx,
                                             double& fy, double& fz) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    ordered_sum3_acc(base, idx, e0, e1, ax, ay, az);
    fx = ax;
    fy = ay;
    fz = az;
}

}  // namespace ms
