// mpm_kernels.hpp -- host-side launchers of the gfx950 kernels (mpm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.hpp"

namespace mpm {

// particles
void launch_set_column(PView P, int col0, int ncomp, const float* aos, const int* slot_of, int n, hipStream_t s);
void launch_get_column(PView P, int col0, int ncomp, float* aos, const int* slot_of, int n, hipStream_t s);
void launch_fill_column(PView P, int col, float value, int n, hipStream_t s);
void launch_identity_columns(PView P, int col0, int n, hipStream_t s);
void launch_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_of_old, int* orig_of_new,
                    int* slot_of, hipStream_t s);
void launch_gather_slots(const int* slot_of, const int* orig, int n, int* slots, hipStream_t s);

// reductions over a slot list: bbox (ordered uint), sum m v and sum m (double)
void launch_body_reduce(PView P, const int* slots, int base, int n, unsigned int* bbox6, double* mom4, hipStream_t s);

// binning / counting sort
void launch_bin(PView P, const int* slots, int base, int n, const GridView& g, int* key, int* rank, int* count, int* err,
                hipStream_t s);
void launch_scatter(const int* slots, int base, const int* key, const int* rank, const int* offset, int n, int* perm,
                    hipStream_t s);
void launch_make_chunks(const int* count, const int* offset, const int* chunk_off, int NT, int4* chunks, hipStream_t s);
void launch_chunk_counts(const int* count, int NT, int* nch, hipStream_t s);
// exclusive scan of n ints; *total (device) receives the sum. tmp needs scan_tmp_ints(n) ints.
int64_t scan_tmp_ints(int n);
void scan_exclusive(const int* in, int* out, int n, int* total, int* tmp, hipStream_t s);

// grid <-> reference dense order
void launch_to_dense_f(const float* src, int ncomp, float* dst, const GridView& g, hipStream_t s);
void launch_from_dense_f(const float* src, int ncomp, float* dst, const GridView& g, hipStream_t s);
void launch_to_dense_i8(const int8_t* src, int8_t* dst, const GridView& g, hipStream_t s);
void launch_from_dense_i8(const int8_t* src, int8_t* dst, const GridView& g, hipStream_t s);

// P2G and friends
void launch_p2g_mass(PView P, const GridView& g, float* mass, int* err, hipStream_t s);
void launch_mass_fix(float* mass, const GridView& g, int* err, hipStream_t s);
void launch_p2g_vel(PView P, const GridView& g, const float* mass, float* vel, int* err, hipStream_t s);
void launch_check_vel(const float* vel, float* prev, const GridView& g, int* err, hipStream_t s);
void launch_volumes(PView P, const GridView& g, const float* mass, int* err, hipStream_t s);

// forces (internal only, into acc) and node epilogues
void launch_forces(PView P, const GridView& g, float* acc, int* err, hipStream_t s);
void launch_node_forces(const GridView& g, const float* mass, const float* acc, const float* grav, int ngrav, float* out,
                        hipStream_t s);
void launch_node_explicit(const GridView& g, const float* mass, const float* vel, const float* acc, const float* grav, int ngrav,
                          float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc, CRScalars* sc,
                          double* partials, hipStream_t s);
void launch_collision_velocities(const GridView& g, const int8_t* flags, const ColliderSet* cols, float* vc,
                                 hipStream_t s);

// HVP: out += -sum_p V dP(dF(d)) F^T grad w   (d projected first when project = true)
void launch_hvp(PView P, const GridView& g, const float* d, bool project, const int8_t* flags, const ColliderSet* cols,
                float* out, const CRScalars* sc, int* err, hipStream_t s);
// diagonal of the force differential (dForceidXi), raw (before -dt^2 and +m)
void launch_diag(PView P, const GridView& g, float* out, int* err, hipStream_t s);

// solver node kernels
void launch_rhs(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* accvc,
                const float* draw, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                double* partials, hipStream_t s);
void launch_cr_setup1(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* acc, const float* x, const float* b, const float* diag, float dt, float* z, float* p,
                      const CRScalars* sc, hipStream_t s);
void launch_cr_setup2(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* acc, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                      double* partials, hipStream_t s);
void launch_cr_update_x(const GridView& g, const int8_t* flags, const ColliderSet* cols, const float* p,
                        const float* Ap, const float* diag, float* x, float* z, CRScalars* sc, double* partials,
                        double* hist, int hist_cap, hipStream_t s);
void launch_cr_ar(const GridView& g, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* acc,
                  const float* z, float dt, float* Ar, CRScalars* sc, double* partials, hipStream_t s);
void launch_cr_update_p(const GridView& g, const float* z, const float* Ar, const float* diag, float* p, float* Ap,
                        CRScalars* sc, double* partials, hipStream_t s);
void launch_residual_vector(const GridView& g, const float* z, const float* diag, float* r, hipStream_t s);
void launch_finalize(const GridView& g, const float* x, const float* vc, float* vel, const CRScalars* sc,
                     hipStream_t s);

// G2P: mode 0 = velocities only, 1 = F only, 2 = full step (v, F, plasticity, advection),
// 3 = full step without advection
void launch_g2p(PView P, const GridView& g, const float* vel, const float* prev, int mode, float dt, float thc,
                float ths, const ColliderSet* cols, int* err, hipStream_t s);
void launch_plasticity(PView P, const int* slots, int base, int n, float thc, float ths, int* err, hipStream_t s);
// ballistic particles: mode bit 1 = velocity update (Sim.cpp:64-73), bit 2 = advection
void launch_ballistic(PView P, const int* slots, int n, const float* grav, int ngrav, float dt, const ColliderSet* cols,
                      int mode, hipStream_t s);
// advection of a slot list (Sim.cpp:123-131)
void launch_advect(PView P, const int* slots, int n, float dt, const ColliderSet* cols, hipStream_t s);
// model probes: mode 0 energy (out n), 1 stress (out 9n), 2 stress differential of dF (9n in, 9n out)
void launch_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out, hipStream_t s);
void launch_check_pvel(PView P, int n, int* err, hipStream_t s);

}  // namespace mpm
