/*
 * nsm_oracle.h -- CPU FP64 restatement of NimbleSM's explicit internal-force
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * product is compared against (and the timed CPU baseline).  Nothing in the
 * product (nimblesm_amd/) links, loads or calls it.
 *
 * Every function follows the reference's operation order so that results are
 * bit-identical to a -ffp-contract=off build of the reference sources
 * (citations are to files of NimbleSM relative to its repository root).
 *
 * Layout conventions (identical to the reference):
 *   node vectors      [N][3]  double, AoS
 *   element conn      [E][8]  int32, 0-based, blocks concatenated in
 *                             ascending block-id order
 *   F per element     [8 ipt][9]  order XX,YY,ZZ,XY,YZ,ZX,YX,ZY,XZ
 *                             (src/nimble_utils.h:101-109)
 *   sigma per element [8 ipt][6]  order XX,YY,ZZ,XY,YZ,ZX
 *                             (src/nimble_utils.h:91-100)
 */
#ifndef NSM_ORACLE_H
#define NSM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element tables (src/nimble_element.cc:55-171) ---- */
void   orc_hex_tables(double int_pts[24], double N[64], double dN[192]);

/* ---- 3x3 kernels (src/nimble_utils.h) ---- */
double orc_invert3x3(const double m[9], double inv[9], int *status);      /* :1229-1268, row-major m[i][j] */
double orc_invert_full33(const double m[9], double inv[9]);               /* :523-551, K_F_* storage       */
double orc_cos_of_acos_div3(double x);                                     /* :650-665                      */
void   orc_eigen_sym33_nonunit(const double t[6], double eval[3],
                               double v0[3], double v1[3], double v2[3]); /* :667-857                      */
void   orc_polar_decomp(const double F[9], double V[6], double R[9]);     /* :859-908                      */

/* ---- material (src/nimble_material.cc:252-310) ---- */
void   orc_neohookean_stress(double K, double G, const double F[9], double sig[6]);

/* ---- element kernels, Kokkos form x = X + u (src/nimble_element.h) ---- */
int    orc_hex_deformation_gradients(const double X[24], const double u[24], double F[72]);  /* :430-502 */
int    orc_hex_nodal_forces(const double X[24], const double u[24], const double sig[48],
                            double f[24]);                                                  /* :539-625 */
int    orc_hex_lumped_mass(double rho, const double X[24], double m[8]);                    /* .cc:173-193, .h:266-309 */
double orc_hex_char_length(const double x[24]);                                             /* .cc:221-260 */

/* ---- batched element / material seams over gathered arrays ---- */
int    orc_batch_deformation_gradients(long E, const double *Xg, const double *ug, double *F);
void   orc_batch_stress(long P, double K, double G, const double *F, double *sig);
int    orc_batch_nodal_forces(long E, const double *Xg, const double *ug, const double *sig, double *f);

/* ---- model level ---- */
typedef struct orc_mesh {
  long          num_nodes;
  long          num_elem;       /* total over blocks                         */
  int           num_blocks;
  const double *X;              /* [N][3]                                    */
  const int    *conn;           /* [E][8]                                    */
  const long   *blk_off;        /* [num_blocks+1] element offsets            */
  const double *rho, *K, *G;    /* [num_blocks] material per block           */
} orc_mesh;

/* BC entries in the order the reference applies them: one entry per
 * (boundary condition, node of its node set).  kind: 0 initial velocity,
 * 1 prescribed velocity, 2 prescribed displacement
 * (src/nimble_boundary_condition_manager.h:93-217). */
typedef struct orc_bcs {
  long          count;
  const int    *node;
  const int    *coord;
  const int    *kind;
  const double *value;
} orc_bcs;

typedef struct orc_state {
  double *u, *v, *a, *f;        /* [N][3] each                                */
  double *m;                    /* [N] lumped mass                            */
  double *Fip, *sip;            /* [E][72], [E][48]; may be NULL              */
  double  time_current;
} orc_state;

void   orc_model_lumped_mass(const orc_mesh *mesh, double *m);                 /* src/nimble_model_data.cc:495-530 */
double orc_model_critical_dt(const orc_mesh *mesh, const double *u);           /* src/nimble_block_base.cc:52-84   */
int    orc_model_internal_force(const orc_mesh *mesh, const double *u, double *f,
                                double *Fip, double *sip);                    /* src/nimble_model_data.cc:620-667 */
/* OpenMP variant: element-parallel forces into per-element slots followed by
 * a node gather in ascending element order; bit-identical to the serial one. */
int    orc_model_internal_force_omp(const orc_mesh *mesh, const double *u, double *f,
                                    double *Fip, double *sip, double *slots, const long *node_ptr,
                                    const long *node_ent);
/* node -> (8*element + local node) incidence, sorted by element, as CSR */
void   orc_build_node_csr(const orc_mesh *mesh, long *node_ptr /*N+1*/, long *node_ent /*8E*/);

void   orc_apply_initial_conditions(const orc_bcs *bcs, double *v);
void   orc_apply_kinematic_bc(const orc_bcs *bcs, double t_cur, double t_prev, const double *u, double *v);

/* Steps [step_begin, step_end) of the central-difference loop
 * (src/integrators/explicit_time_integrator.cc:177-278).  reaction, if not
 * NULL, receives -sum_{reaction_nodes} f[node][reaction_coord] per step
 * (index = step).  threads <= 1 selects the serial internal force. */
int    orc_explicit_steps(const orc_mesh *mesh, const orc_bcs *bcs, orc_state *st,
                          int step_begin, int step_end, int num_load_steps, int output_frequency,
                          double dt_user, const int *reaction_nodes, long num_reaction_nodes,
                          int reaction_coord, double *reaction, int threads);

/* ---- synthetic jittered cube (uniform n^3 HEX8 mesh, SplitMix64 jitter) ---- */
unsigned long long orc_splitmix64(unsigned long long x);
void   orc_jittered_cube(int n, unsigned long long seed, double jitter_frac, double *X, int *conn);

#ifdef __cplusplus
}
#endif
#endif
