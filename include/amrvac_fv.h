/*
 * amrvac_fv.h -- C-ABI of the MI355X-native finite-volume block update.
 *
 * Drop-in boundary: the reference's advect1 (src/mod_advance.t:611-666) and
 * the routines it and setdt reach on the hot path.  A Fortran host binds
 * these entry points through ISO_C_BINDING; every
 * argument is a plain C scalar, pointer or size.  All blocks of one level
 * live in HBM as one batched array per state slot (0 = ps, 1 = ps1,
 * 2 = ps2, 3 = ps3, src/mod_physicaldata.t:111-122).
 *
 * Conventions
 *  - Return value 0 on success, a negative AMRVAC_FV_E* code otherwise; the
 *    host maps a nonzero return onto its mpistop (src/mod_comm_lib.t:207-218).
 *    amrvac_fv_last_error() gives the message.
 *  - Codes for flux schemes, limiters and boundary types are the reference's
 *    own integer parameters (src/mod_global_parameters.t:457-469, :541-550;
 *    src/limiter/mod_limiter.t:15-40).
 *  - Host arrays of block state use the reference layout w(ixG^S,1:nw) per
 *    block (Fortran order: x fastest, then y, z, then the variable), blocks
 *    concatenated in the order given to amrvac_fv_set_mesh.
 *  - Never call from inside an OpenMP parallel region; calls on one context
 *    are serialised on that context's HIP stream.
 */
#ifndef AMRVAC_FV_H
#define AMRVAC_FV_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMRVAC_FV_OK 0
#define AMRVAC_FV_EINVAL (-1)   /* bad argument or unsupported option */
#define AMRVAC_FV_EHIP (-2)     /* HIP runtime error */
#define AMRVAC_FV_ESTATE (-3)   /* call out of order (e.g. no mesh yet) */
#define AMRVAC_FV_ENOMEM (-4)   /* device allocation failed */

/* precision of reconstruction / Riemann / wave speeds; the conserved state,
   flux-divergence accumulation, pressure recovery and dt stay FP64 */
#define AMRVAC_FV_PREC_FP64 0
#define AMRVAC_FV_PREC_MIXED 1

typedef struct amrvac_fv_config {
  int ndim;            /* 2 or 3 (compile-time ndim of the reference) */
  int ndir;            /* vector components, ndim <= ndir <= 3 */
  int physics;         /* 1 = hd (src/hd/mod_hd_phys.t), 2 = mhd (src/mhd/mod_mhd_phys.t) */
  double gamma;        /* hd_gamma / mhd_gamma */
  int divb;            /* typedivbfix: 0 none, 1 glm, 4 linde, 8 ct */
  double glm_alpha;    /* mhd_glm_alpha (0.5) */
  int glm_extended;    /* mhd_glm_extended (1) */
  double divbdiff;     /* Linde coefficient (0.8) */
  int limiter;         /* minmod 1 .. cada3 9, weno5 16 */
  int boundspeed;      /* typeboundspeed: 1 Einfeldt, 2 cmaxmean, 3 cmaxleftright */
  double tvdlfeps;     /* 1.0 */
  double cada3_radius; /* 0.1 */
  double small_pressure, small_density; /* 0.0 */
  int check_small_values;               /* 1 */
  int nghost;          /* 2, or 3 for weno5 */
  int block_nx[3];     /* interior cells per block; block_nx(1) = block_nx(2)
                          = 16, or 8 or 10 in 3-D (FP64) */
  int precision;       /* AMRVAC_FV_PREC_* */
  int type_ct;         /* with divb = 8 (constrained transport on a
                          uniform periodic 2-D or 3-D level, ndir = 3,
                          16-cell blocks): 1 average, 2 uct_contact,
                          3 uct_hll (type_ct,
                          src/mhd/mod_mhd_phys.t:766-781) */
  int fix_small_values; /* 0 or 1 (fix_small_values, methodlist) with the
                          default small_values_method 'error': every cell a
                          stage converts or updates is checked against
                          small_density / small_pressure
                          (src/mod_finite_volume.t:325-333) and a failure
                          raises the crash flag; the HD cmaxmean bound
                          speeds clip the pressure instead
                          (src/hd/mod_hd_phys.t:1040-1045).  Not with
                          constrained transport. */
} amrvac_fv_config;

typedef struct amrvac_fv_ctx amrvac_fv_ctx;

/* Replaces: the physics/method set-up of hd_phys_init / mhd_phys_init and
   read_par_files (src/hd/mod_hd_phys.t:159-370, src/mhd/mod_mhd_phys.t:
   296-949, src/io/mod_input_output.t:292-1080) as far as the block update
   reads it.  Selects the current HIP device's kernels; allocates nothing on
   the device yet. */
int amrvac_fv_init(const amrvac_fv_config* cfg, amrvac_fv_ctx** ctx);
int amrvac_fv_finalize(amrvac_fv_ctx* ctx);

/* HIP stream (hipStream_t passed as void*) for all work of this context;
   NULL selects the context's own stream. */
int amrvac_fv_set_stream(amrvac_fv_ctx* ctx, void* stream);

/* Replaces: the level's connectivity after build_connectivity
   (src/mod_functions_connectivity.t:55) and alloc_node geometry
   (src/amr/mod_amr_solution_node.t:60-186) for one level.
   dx[ndim]: cell size.  neighbor[nb*3^ndim]: same-level neighbour block
   for direction n = (i1+1)+3(i2+1)+9(i3+1), negative = physical boundary.
   xlohi[nb*ndim*2]: block extents.  bc[2*ndim]: boundary type per side.
   Allocates the four state slots (interior cells only, FP64). */
int amrvac_fv_set_mesh(amrvac_fv_ctx* ctx, int nblocks, const double* dx,
                       const int* neighbor, const double* xlohi, const int* bc);

/* Replaces: the connectivity of a block-structured AMR mesh after
   build_connectivity (src/mod_functions_connectivity.t:55-300) and the
   level geometry of alloc_node (src/amr/mod_amr_solution_node.t:150-186).
   The leaves of all levels form one batched array.  dx1[ndim]: level-1
   cell size (level l has dx1/2^(l-1)); level[nb] (1..8);
   ntype[nb*3^ndim]: neighbour type per direction (1 boundary, 2 coarse,
   3 sibling, 4 fine; the centre entry is ignored); nbr[nb*3^ndim]:
   sibling or coarse neighbour leaf (the block itself at the centre);
   nchild[nb*3^ndim*2^ndim]: the children of a fine neighbour (child
   index sum_d (ic_d-1)*2^(d-1)); icpos[nb*ndim]: position within the
   parent (1 or 2); xlohi[nb*ndim*2]: block extents.  Ghost cells across
   level jumps follow getbc (restriction, prolongation with
   interpolation_linear, src/mod_ghostcells_update.t:350-1525) and are kept
   in device-side virtual blocks. */
int amrvac_fv_set_mesh_amr(amrvac_fv_ctx* ctx, int nblocks, const double* dx1, const int* level,
                           const int* ntype, const int* nbr, const int* nchild, const int* icpos,
                           const double* xlohi, const int* bc);

/* Replaces: getbc (src/mod_ghostcells_update.t:350-566) of one slot on an
   AMR mesh (after upload or regrid); no-op on a uniform level, whose ghost
   cells are read from the neighbours by the sweep.  advect1 calls it for
   psb itself. */
int amrvac_fv_getbc(amrvac_fv_ctx* ctx, int slot);

/* Replaces: errest with the Löhner estimator (src/amr/mod_errest.t:10-199,
   lohner_grid; Cartesian, no B0, no log variables) on every leaf of slot
   `slot` (ghost cells as getbc leaves them).  nwt variables iws[] with
   weights[] (w_refine_weight, zero weights left out, ascending iw).
   flags[nb]: bit 0 = some cell has error >= threshold (refine candidate),
   bit 1 = every cell has error < threshold and <= coarsen_threshold
   (derefine_ratio*threshold; coarsen candidate).  The level limits and
   proper nesting stay with the host.  The copy of the flags also fetches
   the small-value flag, so amrvac_fv_crash right after it does not wait on
   the device again. */
int amrvac_fv_errest(amrvac_fv_ctx* ctx, int slot, int nwt, const int* iws,
                     const double* weights, double wavefilter, double threshold,
                     double coarsen_threshold, unsigned char* flags);

/* Replaces: the data movement of amr_coarsen_refine
   (src/amr/mod_coarsen_refine.t:15-173): coarsen_grid
   (src/amr/mod_coarsen.t:12-79) and prolong_2nd (src/amr/mod_refine.t:
   43-167), then build_connectivity and getbc of the new leaves.
   plan[nblocks*(2+2^ndim)] per new leaf: kind 0 copy (old block), 1
   coarsen (the 2^ndim old children in child order), 2 refine (old parent,
   and last the child position sum_d (ic_d-1) 2^d).  The remaining
   arguments describe the new leaves as for amrvac_fv_set_mesh_amr.  Slot 0
   is remapped; slots 1 to 3 are reallocated without data. */
int amrvac_fv_regrid(amrvac_fv_ctx* ctx, int nblocks, const int* plan, const double* dx1,
                     const int* level, const int* ntype, const int* nbr, const int* nchild,
                     const int* icpos, const double* xlohi, const int* bc);

/* Replaces: fix_conserve_at_step (src/mod_advance.t:63-111): when on,
   advect1 on an AMR mesh stores the block-boundary fluxes (store_flux,
   src/mod_fix_conserve.t:597-648) and corrects the coarse cells next to
   finer blocks (fix_conserve, :465-595) before the ghost-cell update. */
int amrvac_fv_set_fix_conserve(amrvac_fv_ctx* ctx, int on);

/* Host <-> device copies of one slot in the reference layout (ghost cells
   included).  upload reads the interior; download fills the ghost cells as
   getbc would (same-level copies and physical boundaries,
   src/mod_ghostcells_update.t:350-566) with a device kernel.
   With constrained transport (divb = 8) each block holds nw + ndim
   variables: the cell variables, then the face field ws(ix, idim) stored
   at the index of the cell whose upper idim face it sits on; combine, copy
   and advect1 carry the face field with the cells, the state-array entry
   points (amrvac_fv_stage, amrvac_fv_slot_pointer) expose the cells only. */
int amrvac_fv_upload(amrvac_fv_ctx* ctx, int slot, const double* w_host);
int amrvac_fv_download(amrvac_fv_ctx* ctx, int slot, double* w_host);

/* Replaces: the crash flag of the reference (src/mod_global_parameters.t:
   646) as the time loop reads it after advance (src/amrvac.t:311-318).
   *crash_out = 1 when a small-value check of any stage, or of the dt
   reduction, since the flag was last cleared would have set crash
   (hd_get_pthermal with check_small_values, src/hd/mod_hd_phys.t:
   1046-1062; the fix_small_values checks of the config).  The flag rides
   the dt copy of amrvac_fv_getdt: read right after it, no synchronisation
   is needed; after a stage the call waits for the stream.  clear != 0
   resets it. */
int amrvac_fv_crash(amrvac_fv_ctx* ctx, int clear, int* crash_out);

/* Replaces: getdt_courant, maxsum (src/mod_dt.t:160-287) over every block,
   with the min over blocks (the MPI_ALLREDUCE of :81-86 on one rank).
   dtmin_out = min_blocks courantpar / max_cells sum_d cmax_d/dx_d;
   cmax_out = max over cells and directions of cmax (cmax_global for GLM,
   :128-129). */
int amrvac_fv_getdt(amrvac_fv_ctx* ctx, int slot, double courantpar,
                    double* dtmin_out, double* cmax_out);

/* Replaces: advect1(method, dtfactor, idim^LIM, qtC, psa, qt, psb)
   (src/mod_advance.t:611-666) for idimmin=1, idimmax=ndim on one level:
   finite_volume (fs_hll/fs_hllc/fs_hlld/fs_tvdlf) or hancock (fs_hancock) for every
   block from slot psa into slot psb with qdt = dtfactor*dt, followed by the
   ghost-cell update of psb (on a uniform level ghost cells are read from
   the neighbours by the next sweep, so nothing is stored).
   Fused RK combination (the whole-array loops of src/mod_advance.t:251-267):
     comb_slot < 0                     psb holds the initial wnew;
     comb_slot >= 0, comb_b == 0       psb = ps[comb_slot] first (ps1 = ps);
     comb_slot >= 0, comb_b != 0       psb = comb_a*ps[comb_slot] + comb_b*ps[psa].
   cmax_global is the GLM speed c_h.  *crash_out (may be NULL) is set to 1
   when a small-value check of the reference would set crash. */
int amrvac_fv_advect1(amrvac_fv_ctx* ctx, int method, double dtfactor,
                      double dt, double cmax_global, int psa, int psb,
                      int comb_slot, double comb_a, double comb_b,
                      int* crash_out);

/* State-array forms for a uniform level, for drivers that keep their states
   in their own device memory (arrays in the slot layout
   [nb][nw][bnz][16][16], ordered on the context's stream):
     amrvac_fv_stage: u_out = init + qdt L(u_in) as advect1 with
       init = comb_a*base + comb_b*u_in (base != NULL) or u_in (base == NULL);
       u_out must not alias u_in (base may alias u_out).
     amrvac_fv_getdt_state: amrvac_fv_getdt of the state u. */
int amrvac_fv_stage(amrvac_fv_ctx* ctx, int method, double dtfactor, double dt,
                    double cmax_global, const double* u_in, const double* base,
                    double comb_a, double comb_b, double* u_out);
int amrvac_fv_getdt_state(amrvac_fv_ctx* ctx, const double* u, double courantpar,
                          double* dtmin_out, double* cmax_out);

/* amrvac_fv_getdt_state without any allocation (no counterpart in the
   reference): for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Its per-block partial maxima live in a file-scope device array of the
   library (512 KiB per GPU, held until the process ends) instead of the
   context's buffers; at most 32768 blocks (a larger level is rejected).
   All calls on one GPU share that array: calls that overlap on one GPU
   race. */
int amrvac_fv_getdt_state_noalloc(amrvac_fv_ctx* ctx, const double* u, double courantpar,
                                  double* dtmin_out, double* cmax_out);

/* Adjoint of one stage (no counterpart in the reference, which is not
   differentiable; the reverse of finite_volume's flux divergence and
   sources, src/mod_finite_volume.t:116-339, on a uniform periodic level with
   Einfeldt bounds and minmod / vanleer / cada3: hd with hll / hllc / tvdlf,
   or ideal mhd (ndir 3) with hll / tvdlf, without divergence control or
   with GLM (typedivbfix 'glm': mhd_modify_wLR, the psi / B_n fluxes and the
   extended sources, src/mhd/mod_mhd_phys.t:6326-6377, 5472-5531)):
     lam_in = coef*lam_out + ca*add_a + cb*add_b
              + (d (qdt L(u)) / d u)^T lam_out  at u = u_in,
   qdt = dtfactor*dt; cmax_global is the GLM speed c_h of the step (held
   fixed, as dt is).  With GLM the stage's psi decay exp(-qdt c_h alpha /
   min dx) is a diagonal factor on the psi row that the caller applies to
   lam_out (and to the add terms of other stages) beforehand.  add_a /
   add_b may be NULL.  All arrays are device arrays in the slot layout;
   lam_in must not alias lam_out or u_in (it may alias add_a or add_b).
   Limiter and Riemann branches are differentiated along the branch the
   forward state takes; the result is deterministic. */
int amrvac_fv_adjoint1(amrvac_fv_ctx* ctx, int method, double dtfactor, double dt,
                       double cmax_global, const double* u_in, const double* lam_out,
                       double coef, const double* add_a, double ca, const double* add_b,
                       double cb, double* lam_in);

/* ps[dst] = ps[src] (interior cells) */
int amrvac_fv_copy(amrvac_fv_ctx* ctx, int dst, int src);

/* Replaces: the state updates of advect between two advect1 calls
   (src/mod_advance.t:113-133 RK2_alfa / ssprk2, 269-291 RK3_BT, 485-507
   rk4, 515-557 ssprk5), e.g. ps = (1/3)*(-ps+ps1+2*ps2+ps3).  ps[dst] is
   set, per interior cell and variable of every leaf, to the postfix program
   ops[0..nops) over the slots: 0..3 push ps, ps1, ps2, ps3; 8 pushes the
   next of consts[0..nconst); 9 add, 10 subtract, 11 multiply, 12 divide,
   13 negate (one IEEE rounding each, no contraction).  At most 32 ops, 16
   constants, 8 deep; dst may be one of the operands. */
int amrvac_fv_combine(amrvac_fv_ctx* ctx, int dst_slot, int nops, const int* ops,
                      int nconst, const double* consts);

/* Replaces: get_volume_average for power 1 and 2
   (src/io/mod_input_output.t:2694-2729); modes[2*nw] */
int amrvac_fv_volume_average(amrvac_fv_ctx* ctx, int slot, double* modes);

/* Device pointer and byte size of a slot (interior-only layout
   [nb][nw][bnz][bny][bnx]), for zero-copy use by PyTorch on the same
   stream. */
int amrvac_fv_slot_pointer(amrvac_fv_ctx* ctx, int slot, void** dptr,
                           long long* nbytes);

/* Device-to-device copy between a slot (interior-only layout) and a device
   array of the same size: to_slot != 0 copies dptr into the slot, else the
   slot into dptr; ordered on the context's stream.  For checkpoints and
   adjoint drivers that keep states in their own device memory.  Rejected
   with constrained transport, whose slots also hold the face field. */
int amrvac_fv_slot_copy(amrvac_fv_ctx* ctx, int slot, void* dptr, int to_slot);

/* Wait for the context's stream. */
int amrvac_fv_sync(amrvac_fv_ctx* ctx);

/* Average duration (ms) of the sweep kernel launches since the last reset,
   measured with HIP events on the context's stream; count = launches. */
int amrvac_fv_sweep_timing(amrvac_fv_ctx* ctx, int reset, double* avg_ms,
                           int* count);

/* Message of the last failing call on this thread. */
const char* amrvac_fv_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
