// selm_forces.hpp -- conservative force terms shared by the bond kernel and
// the auxiliary workgroups of the plane FFT launches.
#pragma once

#include "selm_desc.hpp"

namespace selm {

// LAMMPS bond_style harmonic on bead b (E = K (r - r0)^2, F = -2K (r - r0)
// r_hat, minimum image per dimension), summed in the bead's CSR bond order
// into f0, f1, f2.
__device__ __forceinline__ void bond_terms(const RepDev &R, int b, double &f0, double &f1, double &f2) {
    for (int q = R.adj_ptr[b]; q < R.adj_ptr[b + 1]; q++) {
        const int i1 = R.adj_i1[q], i2 = R.adj_i2[q];
        double del[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double dd = R.x[3 * i1 + d] - R.x[3 * i2 + d];
            if (fabs(dd) > 0.5 * R.Lb[d]) dd += (dd < 0.0) ? R.Lb[d] : -R.Lb[d];
            del[d] = dd;
        }
        const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        const double rr = sqrt(rsq);
        const double dr = rr - R.adj_r0[q];
        const double rk = R.adj_K[q] * dr;
        const double fbond = (rr > 0.0) ? -2.0 * rk / rr : 0.0;
        if (b == i1) {
            f0 += del[0] * fbond;
            f1 += del[1] * fbond;
            f2 += del[2] * fbond;
        } else {
            f0 -= del[0] * fbond;
            f1 -= del[1] * fbond;
            f2 -= del[2] * fbond;
        }
    }
}

// fc of bead g of a batch whose replicas have bonds and external forces
// only (no angle or lj/cut terms): the same sums as k_bond_forces.
__device__ __forceinline__ void bond_only_forces(const RepDev *__restrict__ reps, int nrep,
                                                 const long long *__restrict__ start, long long g) {
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    bond_terms(R, b, f0, f1, f2);
    R.fc[3 * b + 0] = f0 + R.fext[3 * b + 0];
    R.fc[3 * b + 1] = f1 + R.fext[3 * b + 1];
    R.fc[3 * b + 2] = f2 + R.fext[3 * b + 2];
}

// Block roles of a launch of `total` items plus `aux` auxiliary blocks.
// Blocks are dealt round-robin to the 8 XCDs; on each XCD the first
// ceil(total/8) block slots take a contiguous range of items (consecutive
// items share that XCD's L2) and the next ceil(aux/8) slots its share of the
// auxiliary blocks.  Returns the item (>= 0), or -1 with *aux_blk set to the
// auxiliary block (or -1 when the block has nothing to do).
__host__ __device__ inline long long plane_grid(long long total, int aux) {
    return 8 * (((total + 7) >> 3) + ((aux + 7) >> 3));
}
__device__ __forceinline__ long long plane_block_role(long long total, int aux, int *aux_blk) {
    const int xcd = (int)(blockIdx.x & 7);
    const long long slot = blockIdx.x >> 3;
    const long long p8 = (total + 7) >> 3;
    *aux_blk = -1;
    if (slot < p8) {
        const long long g = xcd * p8 + slot;
        return g < total ? g : -1;
    }
    const long long a = (long long)xcd * ((aux + 7) >> 3) + (slot - p8);
    if (a < aux) *aux_blk = (int)a;
    return -1;
}

}  // namespace selm
