// mpm_kernels.hpp -- host-side launchers of the gfx950 kernels (mpm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.hpp"

namespace mpm {

// ---- particle store
void launch_set_column(PView P, int col0, int ncomp, const float* aos, const int* slot_of, int n, hipStream_t s);
void launch_get_column(PView P, int col0, int ncomp, float* aos, const int* slot_of, int n, hipStream_t s);
void launch_fill_column(PView P, int col, float value, int n, hipStream_t s);
void launch_identity_columns(PView P, int col0, int n, hipStream_t s);
void launch_permute(PView src, PView dst, const int* new_to_old, int n, const int* orig_of_old, int* orig_of_new,
                    int* slot_of, hipStream_t s);
void launch_gather_slots(const int* slot_of, const int* orig, int n, int* slots, hipStream_t s);
// device-resident caller indices, range-checked on the device (ERR_INDEX, slot -1)
void launch_gather_slots_checked(const int* slot_of, const int* orig, int n, int np, int* slots, int* err,
                                 hipStream_t s);
void launch_fill_int(int* p, int value, int64_t n, hipStream_t s);
void launch_iota(int* p, int64_t n, hipStream_t s);

// ---- body partition (Sim::calculateBodies) on the GPU
// cell hash with cell edge slightly above h; counts need (mask + 1) ints, zeroed
void launch_cc_hash(PView P, int n, float cell, uint32_t mask, int* bucket, int* rank, int* count, hipStream_t s);
// sorted: cell-ordered float4 (x, y, z, slot bits)
void launch_cc_scatter(PView P, const int* bucket, const int* rank, const int* offset, int n, float4* sorted,
                       hipStream_t s);
// links neighbour pairs (two sampled passes, compression with the pass-1 roots kept in snap, a
// vote for the components pass 2 skips, in skip: kCcSkipBytes of device memory)
constexpr int kCcSkipMax = 8;
constexpr size_t kCcSkipBytes = 512;
void launch_cc_link(int n, float h, float cell, uint32_t mask, const int* offset, const int* count,
                    const float4* sorted, int* parent, int* hasNb, int* snap, void* skip, int* links0, int* links1,
                    hipStream_t s);
// isRoot[s] = 1 for the root of a component of >= 2 particles; minOrig[root] = smallest caller index
void launch_cc_roots(int n, int* parent, const int* hasNb, const int* orig_of, int* isRoot, int* minOrig,
                     hipStream_t s);
// labels[s] = number of the body of particle s (bodies numbered by smallest original index),
// -1 for particles without neighbours; *nbodies (device) = the count. mark, markPos: n ints.
void launch_cc_number(int n, const int* parent, const int* hasNb, const int* isRoot, const int* minOrig, int* mark,
                      int* markPos, int* nbodies, int* tmp, int* labels, hipStream_t s);
void launch_check_pvel(PView P, int n, int* err, hipStream_t s);

// ---- per-body reductions: bbox (ordered uint, 6/body), sum m v and sum m (double, 4/body), count
// per body: bbox (ordered-int encoded min/max), FP64 sums of m v and m, particle count; bboxR,
// momR, countR: body_replicas() copies of the same (scratch)
void launch_body_reduce(PView P, const int* slots, const int* labels, int n, int nb, unsigned int* bbox, double* mom,
                        int* count, unsigned int* bboxR, double* momR, int* countR, hipStream_t s);
int body_replicas();

// ---- binning / counting sort into the batch
void launch_tile_body(const BodyGrid* bodies, const int* order, int nb, int NT, int* tileBody, int* blockBody,
                      hipStream_t s);
void launch_bin(PView P, const int* slots, const int* labels, int n, const BodyGrid* bodies, int* key, int* rank,
                int* count, int* err, hipStream_t s);
// binning keys (k_bin) -> the reference's stencil base node, 3 ints per key (-1 when unbinned)
void launch_decode_bin(const int* key, int n, const BodyGrid* bodies, const int* tileBody, int* cells, hipStream_t s);
void launch_scatter(const int* slots, const int* key, const int* rank, const int* offset, int n, int* perm,
                    int8_t* pcode, hipStream_t s);
void launch_chunk_counts(const int* count, int NT, int* nch, hipStream_t s);
void launch_make_chunks(const int* count, const int* offset, const int* chunk_off, const int* tileBody, int NT,
                        int4* chunks, hipStream_t s);
// exclusive scan of n ints; *total (device) receives the sum. tmp needs scan_tmp_ints(n) ints.
int64_t scan_tmp_ints(int64_t n);
void scan_exclusive(const int* in, int* out, int n, int* total, int* tmp, hipStream_t s);

// ---- batch grid <-> reference dense order of one body
void launch_to_dense_f(const float* src, int ncomp, float* dst, const BodyGrid& G, hipStream_t s);
void launch_from_dense_f(const float* src, int ncomp, float* dst, const BodyGrid& G, hipStream_t s);
void launch_to_dense_i8(const int8_t* src, int8_t* dst, const BodyGrid& G, hipStream_t s);
void launch_from_dense_i8(const int8_t* src, int8_t* dst, const BodyGrid& G, hipStream_t s);

// ---- particle-side kernels: write per-chunk staged footprints
void launch_p2g_mass(PView P, const Batch& B, float* stage, int* err, hipStream_t s);
void launch_p2g_vel(PView P, const Batch& B, const float* mass, float* stage, int* err, hipStream_t s);
void launch_volumes(PView P, const Batch& B, const float* mass, int* err, hipStream_t s);
void launch_forces(PView P, const Batch& B, float* stage, int* err, hipStream_t s);
// HVP: staged -sum_p V dP(dF(d)) F^T grad w (d projected first when project = true); sc may be null
void launch_hvp(PView P, const Batch& B, const float* d, bool project, const int8_t* flags, const ColliderSet* cols,
                float* stage, const CRScalars* sc, int* err, hipStream_t s);
// pairs of consecutive chunks of B that fit one workgroup (groups: up to B.nchunks int2, *ngroups
// device; cnt, pos: (B.nchunks + 1) / 2 ints; tmp: scan_tmp_ints of that), and the HVP over them
// (grid B.nchunks; blocks past *ngroups exit): the same staged footprints as launch_hvp
void launch_pair_chunks(const Batch& B, int* cnt, int* pos, int2* groups, int* ngroups, int* tmp, hipStream_t s);
void launch_hvp_pair(PView P, const Batch& B, const int2* groups, const int* ngroups, const float* d, bool project,
                     const int8_t* flags, const ColliderSet* cols, float* stage, const CRScalars* sc, int* err,
                     hipStream_t s);
// raw diagonal of the force differential (dForceidXi)
void launch_diag(PView P, const Batch& B, float* stage, int* err, hipStream_t s);
// G2P: mode 0 = velocities only, 1 = F only, 2 = full step (v, F, plasticity; advection is
// launch_advect)
void launch_g2p(PView P, const Batch& B, const float* vel, const float* prev, int mode, float dt, float thc, float ths,
                const ColliderSet* cols, int* err, hipStream_t s);

// ---- node kernels (one block per 4 tiles; per-body reductions)
void launch_gather_mass(const Batch& B, const float* stage, float* mass, int* err, hipStream_t s);
void launch_gather_vel(const Batch& B, const float* stage, float* vel, float* prev, int* err, hipStream_t s);
void launch_gather3(const Batch& B, const float* stage, float* out, hipStream_t s);
void launch_node_forces(const Batch& B, const float* mass, const float* stage, const float* grav, int ngrav,
                        float* out, hipStream_t s);
// also marks each tile active (some node with mass or a collider flag) in tileActive
// also writes nrm (the collider's unit normal, 3-vector layout) at every node with a flag >= 0
void launch_node_explicit(const Batch& B, const float* mass, const float* vel, const float* stage, const float* grav,
                          int ngrav, float dt, const ColliderSet* cols, float* em, int8_t* flags, float* vc,
                          CRScalars* sc, BatchStatus* st, double* partials, unsigned char* tileActive, float* nrm,
                          hipStream_t s);
// nrm at every node with a flag >= 0, for flags set by the caller
void launch_node_normals(const Batch& B, const int8_t* flags, const ColliderSet* cols, float* nrm, hipStream_t s);
void launch_collision_velocities(const Batch& B, const int8_t* flags, const ColliderSet* cols, float* vc,
                                 hipStream_t s);
void launch_rhs(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stageVc,
                const float* stageDiag, float dt, float tol, float* b, float* x, float* diag, CRScalars* sc,
                BatchStatus* st, double* partials, hipStream_t s);
void launch_cr_setup1(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* stage, const float* x, const float* b, const float* diag, float dt, float* z,
                      float* p, const CRScalars* sc, hipStream_t s);
void launch_cr_setup2(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                      const float* stage, const float* z, const float* diag, float dt, float* Ap, CRScalars* sc,
                      BatchStatus* st, double* partials, hipStream_t s);
void launch_cr_update_x(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* p, const float* Ap,
                        const float* diag, float* x, float* z, CRScalars* sc, BatchStatus* st, double* partials,
                        double* hist, hipStream_t s);
// Ar = A z, beta and the next alpha (the next denominator by expansion, from Ap and diag)
void launch_cr_ar(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols, const float* stage,
                  const float* z, const float* Ap, const float* diag, float dt, float* Ar, CRScalars* sc,
                  BatchStatus* st, double* partials, hipStream_t s);
// p = z + beta p, Ap = Ar + beta Ap, then the x update of the next iteration (as launch_cr_update_x)
void launch_cr_update_px(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* Ar,
                         const float* diag, float* p, float* Ap, float* x, float* z, CRScalars* sc, BatchStatus* st,
                         double* partials, double* hist, hipStream_t s);
// order each tile's particles by (rank among equal stencil bases, base) (tmpA, tmpB: n ints of
// scratch each, n >= particles binned)
void launch_sort_tiles(const int* count, const int* offset, int* perm, const int8_t* pcode, int* tmpA, int* tmpB,
                       int NT, hipStream_t s);
// small bodies: claim the listed bodies still iterating (done = kDoneSmall), then solve each in
// one workgroup (setup and the whole CR loop)
void launch_small_claim(const int* list, int n, CRScalars* sc, BatchStatus* st, hipStream_t s);
void launch_small_cr(PView P, const Batch& B, const int* list, int n, const float* mass, const int8_t* flags,
                     const ColliderSet* cols, float dt, float* stage, const float* b, const float* diag, float* x,
                     float* z, float* p, float* Ap, float* Ar, CRScalars* sc, double* hist, int* err, hipStream_t s);
// ---- flexible conjugate residual (learned preconditioner path)
void launch_fcr_init(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                     const float* stage, const float* x, const float* b, const float* diag, float dt, float* r,
                     int netOn, double ratio, int mdirs, int strikes, CRScalars* sc, BatchStatus* st,
                     double* partials, hipStream_t s);
void launch_fcr_diag_z(const Batch& B, const float* r, const float* diag, float* z, const CRScalars* sc,
                       hipStream_t s);
void launch_fcr_az(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                   const float* stage, const float* z, const float* diag, float dt, float* Az, const float* Q,
                   int64_t vs, CRScalars* sc, BatchStatus* st, double* partials, hipStream_t s);
// P, Q: fcr_m direction vectors each, vs floats apart
void launch_fcr_orth(const Batch& B, const float* z, const float* Az, const float* r, const float* diag, float* P,
                     float* Q, int64_t vs, CRScalars* sc, BatchStatus* st, double* partials, hipStream_t s);
void launch_fcr_update(const Batch& B, const int8_t* flags, const ColliderSet* cols, const float* mass,
                       const float* diag, const float* P, const float* Q, int64_t vs, float* x, float* r,
                       CRScalars* sc, BatchStatus* st, double* partials, double* hist, hipStream_t s);
// ---- learned preconditioner training on the optimal-step residual loss
// 1 - <r, A z>_D^2 / (|r|_D^2 |A z|_D^2) per body
// z = r / diag + scale clamp(raw) on bodies with weight > 0 (raw may be null: z = r / diag)
void launch_lpc_compose_z(const Batch& B, const float* r, const float* diag, const float* raw, const float* scale,
                          const float* weight, float* z, hipStream_t s);
// with the force differential of z staged: u = A z; dots[3 b + {0, 1, 2}] += <r, u>_D, |u|_D^2, |r|_D^2
void launch_lpc_res_dots(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                         const float* stage, const float* z, const float* r, const float* diag, float dt,
                         const float* weight, float* u, double* dots, hipStream_t s);
// y with dL/dz = A y; *loss += the per-body losses
void launch_lpc_res_y(const Batch& B, const float* u, const float* r, const float* diag, const float* weight,
                      const double* dots, float* y, double* loss, hipStream_t s);
// with A y staged: network output gradient g = scale A y (where the output is not clamped)
void launch_lpc_res_grad(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                         const float* stage, const float* y, const float* raw, float dt, const float* scale,
                         const float* weight, __bf16* g, hipStream_t s);
// ---- training on exact corrections, energy-norm loss 1 - (z.r)^2 / ((z.A z) E) per body
// energy[b] += (x - xk) . rk (x: the solve's final iterate; E of sample k)
void launch_lpc_exact_energy(const Batch& B, const float* x, const float* xk, const float* rk, double* energy,
                             hipStream_t s);
// with the force differential of z staged: u = A z; dots[2 b + {0, 1}] += z . r, z . u
void launch_lpc_energy_dots(const Batch& B, const float* mass, const int8_t* flags, const ColliderSet* cols,
                            const float* stage, const float* z, const float* r, float dt, const float* weight,
                            float* u, double* dots, hipStream_t s);
// network output gradient (g may be null) and *loss += the per-body losses
void launch_lpc_energy_grad(const Batch& B, const float* u, const float* r, const float* raw, const float* scale,
                            const float* weight, const double* dots, const double* energy, __bf16* g, double* loss,
                            hipStream_t s);
void launch_residual_vector(const Batch& B, const float* z, const float* diag, float* r, hipStream_t s);
// ---- plain vectors (mpm_conjugate_residuals): *out = a . b in FP64 (fixed order; partials: 512
// doubles), out = a u + c v (v null: a u)
void launch_vec_dot(const float* a, const float* b, int64_t n, double* partials, double* out, hipStream_t s);
void launch_vec_lincomb(float* out, float a, const float* u, float c, const float* v, int64_t n, hipStream_t s);
void launch_finalize(const Batch& B, const float* x, const float* vc, float* vel, const CRScalars* sc,
                     hipStream_t s);

// ---- particle updates outside the grids
// plasticity of every slot whose label is not `skipLabel` (labels null: all slots)
void launch_plasticity(PView P, const int* labels, int skipLabel, int n, float thc, float ths, int* err,
                       hipStream_t s);
// ballistic particles (label -1): mode bit 1 = velocity update (Sim.cpp:64-73), bit 2 = advection
void launch_ballistic(PView P, const int* labels, int n, const float* grav, int ngrav, float dt,
                      const ColliderSet* cols, int mode, hipStream_t s);
// advection (Sim.cpp:123-131) of every slot with label >= 0
void launch_advect(PView P, const int* labels, int n, float dt, const ColliderSet* cols, hipStream_t s);
// model probes: mode 0 energy (out n), 1 stress (out 9n), 2 stress differential of dF (9n in, 9n out)
void launch_probe(PView P, const int* slots, int n, int mode, const float* dF, float* out, hipStream_t s);
// 1 where labels[s] == -1 (for the ballistic compaction)
void launch_is_ballistic(const int* labels, int n, int* flag, hipStream_t s);
// out[offset + pos[s]] = s for ballistic slots
void launch_ballistic_compact(const int* labels, const int* pos, int n, int* out, int offset, hipStream_t s);

}  // namespace mpm
