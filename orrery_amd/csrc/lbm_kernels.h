// Launchers of the D3Q19 / D2Q9 kernels (defined in lbm3d_kernels.hip, lbm2d_kernels.hip),
// called by the host layer with a filled lattice descriptor and the handle's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_device.h"

namespace lbm {

// One lattice step: bulk kernel over all cells, in B steps the side kernel over the side list,
// boundary kernel over the boundary list. `local` is the A step (odd step count), otherwise the
// B step.
template <class T>
hipError_t d3q19_step(const Lattice3<typename T::R>& p, bool local, int block, int cpt, std::uint32_t lds,
                      hipStream_t s);
// The parts of a step, for per-kernel timing and the two-stream schedule (the side kernel
// launches nothing in A steps).
template <class T>
hipError_t d3q19_step_bulk(const Lattice3<typename T::R>& p, bool local, int block, int cpt, std::uint32_t lds,
                           hipStream_t s);
template <class T> hipError_t d3q19_step_side(const Lattice3<typename T::R>& p, bool local, hipStream_t s);
template <class T>
hipError_t d3q19_step_boundary(const Lattice3<typename T::R>& p, bool local, hipStream_t s);
// Advances a lattice's step clock by one (the end of a step of a dithered 16-bit format).
hipError_t d3q19_tick(std::uint32_t* clock, hipStream_t s);
// The whole step in one launch (bulk and boundary nodes in cell order).
template <class T>
hipError_t d3q19_step_fused(const Lattice3<typename T::R>& p, bool local, hipStream_t s);
// mode 0 uniform u0, 1 moment planes (device), 2 u0 + seeded perturbation.
template <class T>
hipError_t d3q19_initialize(const Lattice3<typename T::R>& p, int mode, const double* planes, const double u0[3],
                            std::uint64_t seed, double amplitude, hipStream_t s);
template <class T>
hipError_t d3q19_populations(const Lattice3<typename T::R>& p, int phase, double* out, hipStream_t s);
template <class T>
hipError_t d3q19_moments(const Lattice3<typename T::R>& p, int phase, const double force[3], double* out,
                         hipStream_t s);
// Moments of n listed nodes into out[4][n]; slots[k] is the boundary slot of cells[k] or -1.
template <class T>
hipError_t d3q19_sample(const Lattice3<typename T::R>& p, int phase, const double force[3], const std::uint32_t* cells,
                        const std::int32_t* slots, std::uint32_t n, double* out, hipStream_t s);
// Largest |u| over fluid nodes into *result (partial: reduction_blocks(cells) + reduction_blocks(nb)).
template <class T>
hipError_t d3q19_max_speed(const Lattice3<typename T::R>& p, int phase, const double force[3], double* partial,
                           double* result, hipStream_t s);
// partial: at least reduction_blocks(cells) + reduction_blocks(nb) doubles; out: 1 double.
template <class T>
hipError_t d3q19_mass(const Lattice3<typename T::R>& p, int phase, double* partial, double* out, hipStream_t s);
// partial: 6 * reduction_blocks(nlinks) doubles; out: 6 doubles (force xyz, moment xyz).
template <class T>
hipError_t d3q19_force_sum(const Lattice3<typename T::R>& p, int phase, const std::uint32_t* lcell,
                           const std::uint8_t* ldir, const double* lq, const std::uint32_t* lslot,
                           std::uint32_t nlinks, double* partial, double* out, hipStream_t s);

template <class T>
hipError_t d2q9_step(const Lattice2<typename T::R>& p, bool local, int block, hipStream_t s);
// The same step as two launches (bulk kernel, then boundary kernel).
template <class T>
hipError_t d2q9_step_split(const Lattice2<typename T::R>& p, bool local, int block, hipStream_t s);
template <class T>
hipError_t d2q9_initialize(const Lattice2<typename T::R>& p, int mode, const double* planes, const double u0[2],
                           std::uint64_t seed, double amplitude, hipStream_t s);
template <class T>
hipError_t d2q9_populations(const Lattice2<typename T::R>& p, int phase, double* out, hipStream_t s);
template <class T>
hipError_t d2q9_moments(const Lattice2<typename T::R>& p, int phase, const double force[2], double* out,
                        hipStream_t s);
// Moments of n listed nodes (slot: boundary slot or -1): out = rho[n], ux[n], uy[n].
template <class T>
hipError_t d2q9_sample(const Lattice2<typename T::R>& p, int phase, const double force[2], const std::uint32_t* cells,
                       const std::int32_t* slots, std::uint32_t n, double* out, hipStream_t s);
// Largest |u| over the fluid nodes into *result; partial as for d2q9_mass.
template <class T>
hipError_t d2q9_max_speed(const Lattice2<typename T::R>& p, int phase, const double force[2], double* partial,
                          double* result, hipStream_t s);
template <class T>
hipError_t d2q9_mass(const Lattice2<typename T::R>& p, int phase, double* partial, double* out, hipStream_t s);
// partial: 3 * reduction_blocks(nlinks) doubles; out: 3 doubles (fx, fy, moment).
template <class T>
hipError_t d2q9_force_sum(const Lattice2<typename T::R>& p, int phase, const std::uint32_t* lcell,
                          const std::uint8_t* ldir, const std::uint32_t* lslot, std::uint32_t nlinks,
                          double* partial, double* out, hipStream_t s);

// Blocks of 256 threads the reductions use for n items.
std::uint32_t reduction_blocks(std::uint64_t n);

} // namespace lbm
