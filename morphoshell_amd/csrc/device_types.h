// Types shared by the host context and the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ms {

// Per-shell scalars and status, resident in device memory (one entry per
// shell).  Kernels read them as block-uniform values through the chunk
// tables; the last block of each node launch does the per-step bookkeeping
// (buffer parity, step clock, dial clock, equilibrium verdicts).
struct ShellDev {
    // mechanics (calc_deformation_forces.cpp:93-95, calc_energy_densities.cpp:62-77)
    double nu;
    double gamma;          // bend_stiffening_scale_factor
    // dynamics (calc_non_deformation_forces.cpp:55-57, advance_dynamics.cpp:58-93)
    double dt;
    double eta;            // damping_factor
    double damp_den;       // ref_density * h_min * h_min
    double lim;            // 5e4 * char_force_scale
    double fchar;
    double cchar;
    // dialling clock (main.cpp:286-288, 478-488)
    double dial_last_eq;
    double dial_spacing;
    double dial_K;         // timesteps_per_dial_phase as double
    long long phase_steps;
    int dialling;
    int protocol;          // 1 blend, 3 LCE
    int from_ansatz;
    int gd;                // gradient-descent dynamics
    int has_a0;
    int has_b0;
    // ranges
    int node_begin, num_nodes;
    int tri_begin, num_tris;
    int node_chunk_begin, node_chunk_count;
    int tri_chunk_begin, tri_chunk_count;
    // state
    int par;               // which of the two x/v buffers holds the current state
    int halted;            // 1: skipped by ms_step
    int failed;            // set by a node launch that found |F| >= lim
    int at_eq;
    long long halt_step;   // step-clock value at which the shell halted
    long long fail_step;
    // last equilibrium check
    double max_f;
    double max_v;
    int f_node;
    int v_node;
    int checked;
    int pad_;
};

// A contiguous range of triangles or nodes of one shell, processed by one
// block: {begin, end, shell, unused}.
typedef int4 Chunk;

// Global clock of a context: steps taken and the completion counter used by
// the last-block bookkeeping of node launches.
struct ClockDev {
    long long steps;
    unsigned int done;
    unsigned int pad_;
};

}  // namespace ms
