/*
 * oracle.h -- CPU restatement of TCLB's per-time-step node kernel.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libtclbx.so) never links or calls it.
 *
 * It restates, in plain C, what the reference CPU build (CROSS_CPU, double)
 * computes for the models in scope:
 *   d2q9_SRT       models/flow/d2q9_SRT/Dynamics.c:26-176, Dynamics.R:6-37
 *   d2q9_SRT_drag  the same, plus Ladd momentum-exchange drag on Solid nodes
 *                  in the form of models/flow/d3q27/Dynamics.c.Rt:68-82
 *   d3q19          models/flow/d3q19/Dynamics.c.Rt:23-290, Dynamics.R:1-95,
 *                  src/lib/d3q19.R:24-96
 *   d2q9           models/flow/d2q9/Dynamics.c.Rt, Dynamics.R with the basis of
 *                  src/lib/feq.R and the Zou-He of src/lib/boundary.R:155-207
 *   d2q9_drag      d2q9 plus Drag/Lift on Solid nodes (as d2q9_SRT_drag)
 *   d3q27          models/flow/d3q27/Dynamics.c.Rt, Dynamics.R with the raw
 *                  moment basis of src/lib/feq.R (ortogonal=FALSE), Zou-He of
 *                  src/lib/boundary.R:66-109, Symmetry of :33-50
 *   d2q9_autosym, d2q9_autosym2
 *                  d2q9 with the symmetric access of its autosym builds
 *                  (src/conf.R:373-445, src/LatticeAccess.inc.cpp.Rt:233-470)
 *   d2q9_bc, d2q9_bc_autosym, d2q9_bc_autosym2
 *                  d2q9's bc builds: parameter densities BC[0], BC[1]
 *                  (models/flow/d2q9/Dynamics.R:23-26, Dynamics.c.Rt:28-34,
 *                  49-53,80-87,146-159,254-261)
 *   d3q27_cumulant models/flow/d3q27_cumulant/Dynamics.c.Rt, Dynamics.R, and
 *                  its option builds: SMAG (Smagorinsky, :362-374), IB
 *                  (interpolated bounce-back on cut distances, :180-222), AVG
 *                  (averaged densities, Dynamics.R:66-90, Dynamics.c.Rt:62-99,
 *                  410-414,507-518), in every combination
 *   d3q19_lid      d3q19 plus a MovingWall node type (bounce-back with Ladd's
 *                  wall-velocity term) and its force/torque Globals; the
 *                  reference has no moving wall, so this part is pinned by
 *                  closed forms only (Couette flow, momentum balance)
 * with pull streaming on a periodic torus (src/LatticeAccess.inc.cpp.Rt:
 * 319-347, src/Lattice.cu.Rt:445-453), the Run() node order of
 * src/cuda.cu.Rt:322-327 and Globals on the last step of a segment
 * (src/Lattice.cu.Rt:900-989).
 *
 * Storage emulation (src/LatticeAccess.inc.cpp.Rt:14-35): F64 keeps double
 * arithmetic; the reduced formats compute in float and store (f - w_i) in
 * float / IEEE half / bfloat16 with round-to-nearest-even.
 *
 * Globals are accumulated in long double (x87 80-bit) so the oracle's sums
 * are at least as accurate as any summation order of the device.
 *
 * Pinned in tests/test_oracle_pins.py by closed forms, independent
 * derivations and brute force (each test names what it pins); the cumulant
 * model's Galilean-correction terms by what they are for, the U^2 defect
 * they remove from a longitudinal wave's decay under a mean flow
 * (test_cumulant_galilean_correction_restores_galilean_invariance).  Parity
 * unpinned -- nothing the reference holds fixes these, since its golden
 * data (tests/external) is absent and its generator (R polyAlgebra) cannot
 * run here:
 *   - the floating-point operation order that polyAlgebra's ToC emits for
 *     the generated models (d3q19 CollisionMRT / SetEquilibrum / Zou-He, the
 *     d2q9 MRT, d3q27 and cumulant collisions and Zou-He); the oracle follows
 *     the dense forms in ascending index order, within the FP64 tolerance;
 *   - the reference float build's literal rewrite (Dynamics_sp.c,
 *     src/makefile.main.Rt:383-386), which the reduced formats do not follow.
 */
#ifndef TCLBX_ORACLE_H
#define TCLBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_D2Q9_SRT = 0, ORC_D2Q9_SRT_DRAG = 1, ORC_D3Q19 = 2, ORC_D3Q19_LID = 3, ORC_D2Q9 = 4,
       ORC_D2Q9_DRAG = 5, ORC_D3Q27 = 6, ORC_D3Q27_CUMULANT = 7, ORC_D2Q9_AUTOSYM = 8, ORC_D2Q9_AUTOSYM2 = 9,
       ORC_D3Q27_CUMULANT_SMAG = 10, ORC_D3Q27_CUMULANT_IB = 11, ORC_D3Q27_CUMULANT_IB_SMAG = 12,
       ORC_D3Q27_CUMULANT_AVG = 13, ORC_D3Q27_CUMULANT_AVG_IB = 14, ORC_D3Q27_CUMULANT_AVG_SMAG = 15,
       ORC_D3Q27_CUMULANT_AVG_IB_SMAG = 16, ORC_D2Q9_BC = 17, ORC_D2Q9_BC_AUTOSYM = 18, ORC_D2Q9_BC_AUTOSYM2 = 19 };
enum { ORC_F64 = 0, ORC_F32 = 1, ORC_F32S = 2, ORC_F16S = 3, ORC_BF16S = 4 };

/* d2q9 parameter vector:   [omega, Density]
 * d2q9 zone columns:       [Velocity, Velocity_x, Velocity_y, GravitationX,
 *                           GravitationY, (drag model only) DragInObj, LiftInObj]
 * d3q19 parameter vector:  [omega, InletVelocity, InletDensity, ForceX, ForceY,
 *                           ForceZ, S_e, S_eps, S_q, S_pi, S_m]   (S_* <= 0: the
 *                           reference's fixed rates); d3q19_lid appends
 *                           [MovingWallVelocityX, Y, Z]
 * d3q19 zone columns:      24 <Global>InObj weights in Globals order (28 for
 *                           d3q19_lid).
 * d2q9 (MRT) parameters:   [S2, S3, S4, GravitationX, GravitationY]
 * d2q9 (MRT) zone columns: [VelocityX, VelocityY, Pressure, PressureLossInObj,
 *                           OutletFluxInObj, InletFluxInObj]
 * d3q27 parameters:        [omega, Smag, ForceX, ForceY, ForceZ]
 * d3q27 zone columns:      [Velocity, Pressure, Turbulence, then the six
 *                           <Global>InObj weights in Globals order]
 * d3q27_cumulant params:   [nu, nubuffer, GalileanCorrection, ForceX, ForceY,
 *                           ForceZ, Omega, (d3q27_cumulant_SMAG) Smag]
 * d3q27_cumulant zones:    [Velocity, Pressure, Turbulence, then the nine
 *                           <Global>InObj weights in Globals order]         */

typedef struct orc_lattice orc_lattice;

orc_lattice* orc_create(int model, int nx, int ny, int nz, int storage);
void orc_destroy(orc_lattice* L);
int  orc_q(const orc_lattice* L);
int  orc_nglobals(const orc_lattice* L);
void orc_set_params(orc_lattice* L, const double* p, int n);
void orc_set_zone_table(orc_lattice* L, const double* table, int nzones, int ncols);
void orc_set_flags(orc_lattice* L, const uint16_t* flags);
/* Cut distances of the 26 links of every node, cuts[d*N + n], d = direction
 * - 1 in the d3q27_cumulant order (Lattice::CutsOverwrite of the whole
 * lattice, src/Lattice.cu.Rt:1056-1072); without a call every link reads
 * NO_CUT (src/LatticeAccess.inc.cpp.Rt:279-283). */
void orc_set_cuts(orc_lattice* L, const uint16_t* cuts);
/* The densities that are not streamed, [k][N] in FP64: the AVG builds' 13
 * averages (avgP, varUX, varUY, varUZ, varUXUY, varUXUZ, varUYUZ, avgdxu2,
 * avgdyv2, avgdzw2, avgUX, avgUY, avgUZ; quantities 5 KinE, 6 ReStr,
 * 7 Dissipation, 8 avgU, 9 varU, 10 averageP of orc_get_quantity) or the bc
 * builds' BC[0], BC[1].  resetAverage (src/Lattice.cu.Rt:1360-1365) zeroes
 * the averages. */
void orc_reset_average(orc_lattice* L);
int orc_nextra(const orc_lattice* L);
void orc_get_extra(const orc_lattice* L, double* out);
void orc_set_extra(orc_lattice* L, const double* in);
/* InterpolatedBounceBack alone on the pulled f of one node (FP64): cut[26]
 * of its links, old[27] its populations in the snapshot being read, and
 * Drag/Lateral/Lift added to globals[2], [4], [3] on Body nodes. */
void orc_cum_ib(double f[27], uint16_t nt, const uint16_t cut[26], const double old[27], const double* zone,
                double* globals);
void orc_init(orc_lattice* L);
/* niter steps; if glob_last, Globals of the last step go to globals_out. */
void orc_iterate(orc_lattice* L, int niter, int glob_last, double* globals_out);
/* f as real values (storage -> real + shift), layout [q][z][y][x]. */
void orc_get_f(const orc_lattice* L, double* out);
void orc_set_f(orc_lattice* L, const double* in);
/* quantity: 0 = Rho, 1 = U (3 doubles per node, interleaved), 2 = P,
 * 3 = Fd (d3q27, 3 doubles per node), 4 = Solid (d3q27_cumulant).
 * Evaluated like the reference get<Q> kernels: on pulled (streamed)
 * populations of the current snapshot, without boundary or collision. */
void orc_get_quantity(const orc_lattice* L, int which, double* out);

/* threads of the step loop (OpenMP over (y,z) rows); 1 by default */
void orc_set_threads(int n);

/* storage conversions, exported so tests can pin them */
uint16_t orc_f32_to_f16(float v);
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float v);
float    orc_bf16_to_f32(uint16_t h);
/* stochastic rounding of the 16-bit formats (an engine option): the random
 * word of (node, direction, step, seed), and the two conversions (gfx950's
 * v_cvt_sr_f16_f32 and v_cvt_sr_bf16_f32 restated) */
uint32_t orc_sr_bits(uint32_t node, uint32_t q, uint32_t step, uint32_t seed);
uint16_t orc_f32_to_f16_sr(float x, uint32_t r);
uint16_t orc_f32_to_bf16_sr(float x, uint32_t r);
void orc_set_stochastic(orc_lattice* L, int on, uint32_t seed);

/* single-node kernels exported for closed-form tests (double arithmetic) */
void orc_d2q9_feq(double feq[9], double d, double ux, double uy);
void orc_d3q19_feq(double feq[19], double rho, double jx, double jy, double jz);
void orc_d3q19_collide(double f[19], const double* params);
void orc_d3q27_feq(double feq[27], double rho, double jx, double jy, double jz);
/* one node update of d3q27 (boundary + collision); globals[7] if not NULL */
void orc_d3q27_node(double f[27], uint16_t nt, const double* params, const double* zone, double* globals);
/* the moment basis: M[27*i + j] = moment j of direction i, its order and powers */
void orc_d3q27_basis(double M[27 * 27], int order[27], int p[27 * 3]);
/* one node update of d3q27_cumulant (smag: its SMAG build); globals[10] if not NULL */
void orc_cum_node(double f[27], uint16_t nt, const double* params, const double* zone, double* globals, int smag);

#ifdef __cplusplus
}
#endif
#endif
