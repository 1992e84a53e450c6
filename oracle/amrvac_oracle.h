/*
 * amrvac_oracle.h -- CPU restatement (FP64, plain C) of MPI-AMRVAC's
 * finite-volume block update, used ONLY as test infrastructure: the parity
 * checker for the HIP path and the CPU baseline timed beside it.  Nothing in
 * the product links this library.
 *
 * Layout follows the reference's per-block state w(ixG^S,1:nw)
 * (src/mod_physicaldata.t:5-50): Fortran column-major, x fastest, variable
 * slowest, ghost cells included; one contiguous array per block, blocks
 * concatenated: w[slot][iblock][iw][k][j][i].
 *
 * Codes (flux schemes, limiters, boundary types) are the reference's own
 * integer parameters (src/mod_global_parameters.t:457-469, :541-550;
 * src/limiter/mod_limiter.t:15-40) so that the host can pass them through
 * unchanged.
 */
#ifndef AMRVAC_ORACLE_H
#define AMRVAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* physics */
#define ORC_PHYS_HD 1
#define ORC_PHYS_MHD 2
/* div B control (src/mhd/mod_mhd_phys.t:96-105) */
#define ORC_DIVB_NONE 0
#define ORC_DIVB_GLM 1
#define ORC_DIVB_LINDE 4
#define ORC_DIVB_CT 8
/* type_ct of constrained transport (src/mhd/mod_mhd_phys.t:766-781) */
#define ORC_CT_AVERAGE 1
#define ORC_CT_CONTACT 2
#define ORC_CT_HLL 3
/* flux schemes (src/mod_global_parameters.t:457-469) */
#define ORC_FS_HLL 1
#define ORC_FS_HLLC 2
#define ORC_FS_HLLD 3
#define ORC_FS_TVDLF 5
#define ORC_FS_HANCOCK 8
/* limiters (src/limiter/mod_limiter.t:15-40) */
#define ORC_LIM_MINMOD 1
#define ORC_LIM_WOODWARD 2
#define ORC_LIM_MCBETA 3
#define ORC_LIM_SUPERBEE 4
#define ORC_LIM_VANLEER 5
#define ORC_LIM_ALBADA 6
#define ORC_LIM_KOREN 7
#define ORC_LIM_CADA 8
#define ORC_LIM_CADA3 9
#define ORC_LIM_WENO5 16
/* physical boundary types (src/mod_global_parameters.t:541-550) */
#define ORC_BC_CONT 2
#define ORC_BC_SYMM 3
#define ORC_BC_ASYMM 4
#define ORC_BC_PERIODIC 5

typedef struct orc_params {
  int ndim, ndir;          /* compile-time ndim / ndir of the reference */
  int phys;                /* ORC_PHYS_* */
  double gamma;            /* hd_gamma or mhd_gamma */
  int divb;                /* ORC_DIVB_* (MHD only) */
  double glm_alpha;        /* mhd_glm_alpha, default 0.5 */
  int glm_extended;        /* mhd_glm_extended, default 1 */
  double divbdiff;         /* Linde coefficient, default 0.8 */
  int limiter;             /* ORC_LIM_* */
  int boundspeed;          /* 1 Einfeldt, 2 cmaxmean, 3 cmaxleftright */
  double tvdlfeps;         /* default 1 */
  double cada3_radius;     /* default 0.1 */
  double small_pressure, small_density; /* defaults 0 */
  int check_small_values;  /* default 1 */
  int nghost;              /* 2 (3 for weno5) */
  int block_nx[3];         /* interior cells per block */
  int type_ct;             /* ORC_CT_* when divb = ORC_DIVB_CT */
  int fix_small_values;    /* default 0; small_values_method 'error' */
} orc_params;

typedef struct orc_sim orc_sim;

/* number of conserved variables for these params */
int orc_nw(const orc_params* p);

/*
 * nblocks blocks on one level with cell size dx[ndim].
 * neighbor[ib*3^ndim + n]: same-level neighbour block index for direction
 *   n = (i1+1) + 3*(i2+1) + 9*(i3+1), i_d in {-1,0,1}; negative = physical
 *   boundary (the bc type of that side is then taken from bc[2*d+side]).
 * xlohi[(ib*ndim + d)*2 + {0,1}]: block extent rnode(rpxmin/rpxmax)
 *   (src/amr/mod_amr_solution_node.t:164-167).
 * bc[2*ndim]: boundary type per side (min1,max1,min2,max2,...).
 */
int orc_create(const orc_params* p, int nblocks, const double* dx,
               const int* neighbor, const double* xlohi, const int* bc,
               orc_sim** out);
void orc_destroy(orc_sim* s);

/* whole-slot copies in the reference layout (4 slots: ps .. ps3).  With
   constrained transport (divb = ORC_DIVB_CT; uniform periodic meshes) each
   block holds nw + ndim variables: the cell variables, then the face field
   ws(ix, idim) of the reference (src/mod_physicaldata.t), stored at the
   index of the cell whose upper idim face it sits on; the ghost fill and
   every slot combination carry it like a cell variable. */
int orc_set_w(orc_sim* s, int slot, const double* w);
int orc_get_w(const orc_sim* s, int slot, double* w);

/* cell-centre coordinates of every block in the reference layout
   (src/amr/mod_amr_solution_node.t:180-186), x[ib][d][k][j][i] */
int orc_get_x(const orc_sim* s, double* x);

/* getdt_courant (maxsum) over all blocks of a slot
   (src/mod_dt.t:160-287): returns min dt over blocks and max cmax */
int orc_getdt_courant(orc_sim* s, int slot, double courantpar,
                      double* dtmin_out, double* cmax_out);

/* advect1 (src/mod_advance.t:611-666): for every block,
   finite_volume / hancock from slot psa into slot psb, then getbc(psb).
   qdt = dtfactor*dt; cmax_global is used by GLM. */
int orc_advect1(orc_sim* s, int method, double dtfactor, double dt,
                double cmax_global, int psa, int psb);

/* dst = a*x + b*y over every cell (ghosts included), the RK combinations
   of src/mod_advance.t:251-267 */
int orc_axpby(orc_sim* s, int dst, double a, int x, double b, int y);
int orc_copy(orc_sim* s, int dst, int src);
int orc_combine(orc_sim* s, int dst, int nops, const int* ops, const double* k);

/* getbc (src/mod_ghostcells_update.t:350-566) */
int orc_getbc(orc_sim* s, int slot);

/* volume-weighted mean(w**power) over interiors
   (src/io/mod_input_output.t:2694-2729); modes[2*nw] */
int orc_volume_average(const orc_sim* s, int slot, double* modes);

/* set by any small-value check that the reference would flag */
int orc_crash_flag(const orc_sim* s);

/* unit-test hooks: dwlimiter2 for one cell (a = dwC(i-1), b = dwC(i));
   WENO5 for one face from w6 = w(i-2..i+3); one face of finite_volume from
   nl primitive cells lw[iw*nl + k] around the face (between nl/2-1 and
   nl/2); the cell cmax of a conserved state */
int orc_dwlimiter(const orc_sim* s, int idim, double a, double b, double* ldw, double* rdw);
int orc_weno5_face(const double* w6, double* wl, double* wr);
int orc_face_flux(orc_sim* s, int method, int idim, int nl, const double* lw, double ch,
                  double* fC, double* cmax, double* cmin);
double orc_cell_cmax(const orc_sim* s, const double* wcons, int idim);

/* block-structured AMR: dxb[nb][ndim] cell size per block; ntype/nbr
   [nb][3^ndim] neighbour type (1 boundary, 2 coarse, 3 sibling, 4 fine; 0
   for the block itself) and sibling / coarse leaf; nchild[nb][3^ndim]
   [2^ndim] children of a fine neighbour (child index sum((ic_d-1) << d));
   icpos[nb][ndim] position within the parent (1 or 2).  getbc then follows
   the level-jump path of src/mod_ghostcells_update.t and, when enabled,
   advect1 stores block-boundary fluxes and applies fix_conserve
   (src/mod_fix_conserve.t:465-648). */
int orc_create_amr(const orc_params* p, int nblocks, const double* dxb, const int* ntype,
                   const int* nbr, const int* nchild, const int* icpos, const double* xlohi,
                   const int* bc, orc_sim** out);
void orc_set_fix_conserve(orc_sim* s, int on);

/* regrid on single ghosted blocks: coarsen child ic into the parent's
   quadrant (coarsen_grid_siblings), prolong_2nd child ic from the parent,
   Löhner error estimate of the interior cells (lohner_grid) */
int orc_coarsen_child(const orc_sim* s, const double* wFi, const int* ic, double* wCo);
int orc_prolong_2nd(const orc_sim* s, const double* wCo, const int* ic, double* wFi);
int orc_lohner_error(const orc_sim* s, const double* w, const double* weight, double wavefilter,
                     double* err);

/* number of OpenMP threads the oracle uses (0 = library default) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
