/*
 * TEST INFRASTRUCTURE ONLY.  FP64 CPU oracle for the Phase discretise-and-solve
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker or the CPU baseline; the
 * product path (phase_amd, libphase_hip.so) never links or calls it.
 *
 * Every function restates a routine of the reference (obrienadam/Phase); the
 * file:line it follows is given beside it, paths relative to the reference's
 * root.  All arithmetic is IEEE binary64 compiled with -ffp-contract=off, so
 * each result is the exact sequence of roundings written here.
 *
 * Conventions: cell/face/node ids are 0-based int32 (reference Index=int,
 * src/Types/Types.h:1-14); vector fields are interleaved (x0,y0,x1,y1,...);
 * per-face boundary types use ORC_BC_* below.  Sign convention of a solve is
 * the reference's: the stored equation A x + rhs = 0 is solved as
 * A x = -rhs (src/2D/Unstructured/FiniteVolume/Equation/FiniteVolumeEquation.tpp:73).
 */
#ifndef PHASE_ORACLE_H
#define PHASE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same values as FiniteVolumeField<T>::BoundaryType
 * (src/2D/Unstructured/FiniteVolume/Field/FiniteVolumeField.h:12). */
enum { ORC_BC_FIXED = 0, ORC_BC_NORMAL_GRADIENT = 1, ORC_BC_SYMMETRY = 2,
       ORC_BC_OUTFLOW = 3, ORC_BC_PARTIAL_SLIP = 4 };

/* ---- mesh: FiniteVolumeGrid2D::init(nodes,cptr,cind) ---------------------- */
void *orc_mesh_create(int nNodes, const double *xy, int nCells, const int *cptr,
                      const int *cind, char *err, int errLen);
void orc_mesh_free(void *mesh);
void orc_mesh_sizes(const void *mesh, int *nFaces, int *nInteriorFaces,
                    int *nBoundaryFaces);
void orc_mesh_faces(const void *mesh, int *lNode, int *rNode, int *lCell,
                    int *rCell, double *centroid, double *sf);
void orc_mesh_cells(const void *mesh, double *volume, double *centroid);
void orc_mesh_links(const void *mesh, int *iptr, int *iface, int *inb,
                    int *bptr, int *bface);

/* ---- padded CSR equation: CrsEquation ------------------------------------- */
void *orc_eqn_create(int nRows, int nnzPerRow);
void *orc_eqn_copy(const void *eqn);
void orc_eqn_free(void *eqn);
int orc_eqn_rank(const void *eqn);
int orc_eqn_nslots(const void *eqn);
void orc_eqn_get(const void *eqn, int *rowPtr, int *colInd, double *vals,
                 double *rhs);
void orc_eqn_add_coeff(void *eqn, int row, int col, double val);
void orc_eqn_set_coeff(void *eqn, int row, int col, double val);
void orc_eqn_add_rhs(void *eqn, int row, double val);
void orc_eqn_plus_eqn(void *lhs, const void *rhs, int sign);
void orc_eqn_minus_vector(void *eqn, const double *v);
void orc_eqn_eq_scalar(void *eqn, double s);
void orc_eqn_scale(void *eqn, double s);
void orc_eqn_relax(void *eqn, double alpha, const double *phi);

/* ---- fv:: operators -------------------------------------------------------- */
/* variant: 0 = laplacian(Scalar gamma, phi, theta)   Laplacian.h:7-47
 *          1 = laplacian(Scalar gamma, phi)          Laplacian.h:49-83
 *          2 = laplacian(field gamma, phi, theta)    Laplacian.h:85-130
 *          3 = laplacian(field gamma, phi)           Laplacian.h:132-167
 * ncomp 2 selects the Vector2D specialisations (Laplacian.cpp:5-120) for
 * variants 0 and 2.  Returns NULL and fills err on an unsupported boundary. */
void *orc_laplacian(const void *mesh, int variant, int ncomp, double gamma,
                    const double *gammaFace, const double *gamma0Face,
                    double theta, const int *bcType, const double *phiFace,
                    const double *phi0Cell, const double *phi0Face, char *err,
                    int errLen);
void *orc_ddt(const void *mesh, int ncomp, double timeStep,
              const double *phi0Cell);
void *orc_div(const void *mesh, int ncomp, const double *uFace,
              const double *u0Face, double theta, const int *bcType,
              const double *phiFace, const double *phi0Cell,
              const double *phi0Face, char *err, int errLen);
void orc_src_div(const void *mesh, const double *uFace, double *out);
void *orc_divc(const void *mesh, int ncomp, const double *uFace, const double *u0Face, double theta,
               const int *bcType, const double *phiFace, const double *phi0Cell, const double *phi0Face, char *err,
               int errLen);
void *orc_dive(const void *mesh, int ncomp, const double *u0Face, const double *u1Face, double theta,
               const int *bcType, const double *phi0Cell, const double *phi1Cell, const double *phi0Face,
               const double *phi1Face, char *err, int errLen);
void *orc_div_grad(const void *mesh, const double *phiUFace, const double *gradU, const int *bcType,
                   const double *uFace, char *err, int errLen);
void *orc_ddt_rho(const void *mesh, int ncomp, double timeStep, double rho, const double *rhoCell,
                  const double *rho0Cell, const double *phi0Cell, const int *cellMask);
void orc_eqn_scale_cells(void *eqn, int nCells, const double *field);
void orc_eqn_remove_cells(void *eqn, int nCells, const int *mask);
void orc_face_to_cell(const void *mesh, const double *faceVec, const double *cellWeight, const double *faceWeight,
                      double *out);
double orc_cicsam_hc(double g, double co);
double orc_cicsam_uq(double g, double co);
void orc_cicsam_weights(const void *mesh, const double *uFace, const double *gammaCell, const double *gradCell,
                        double timeStep, double *beta);
void *orc_cicsam_div(const void *mesh, const double *uFace, const double *beta, double theta, const int *bcType,
                     const double *gammaFace, const double *gamma0Cell, const double *gamma0Face, char *err,
                     int errLen);
void orc_cicsam_momentum_flux(const void *mesh, double rho1, double rho2, const double *uFace,
                              const double *gammaCell, const double *gammaFace, const double *beta, double *rhoU);
void orc_vof_properties(const void *mesh, double rho1, double rho2, double mu1, double mu2,
                        const double *gammaCell, const double *gammaFace, double *rhoCell, double *rhoFace,
                        double *muCell, double *muFace);
void orc_src_src(const void *mesh, int ncomp, const double *phiCell,
                 double *out);
void orc_grad(const void *mesh, const double *phiCell, const double *phiFace,
              double *gradFace, double *gradCell);
void orc_interpolate_faces(const void *mesh, int ncomp, const double *phiCell,
                           const int *bcType, double *phiFace);

/* ---- linear algebra -------------------------------------------------------- */
void orc_spmv(int n, const int *rowPtr, const int *colInd, const double *vals,
              const double *x, double *y);
int orc_ilu0(int n, const int *rowPtr, const int *colInd, double *vals);
int orc_bicgstab_ilu0(int n, const int *rowPtr, const int *colInd,
                      const double *vals, const double *b, double *x,
                      double tol, int maxIters, int *iters, double *relres);
/* OpenMP variant of orc_bicgstab_ilu0 (SpMV, dots and vector updates over
 * nThreads threads; ILU(0) sweeps sequential): CPU timing baseline only. */
int orc_bicgstab_ilu0_omp(int n, const int *rowPtr, const int *colInd, const double *vals, const double *b,
                          double *x, double tol, int maxIters, int nThreads, int *iters, double *relres);
int orc_ilut(int n, const int *rowPtr, const int *colInd, const double *vals,
             double dropTol, int fillFactor, int **outRowPtr, int **outColInd,
             double **outVals);
int orc_bicgstab_ilut(int n, const int *rowPtr, const int *colInd,
                      const double *vals, const double *b, double *x,
                      double tol, int maxIters, double dropTol, int fillFactor,
                      int *iters, double *relres);
void orc_lu_solve(int n, const int *rowPtr, const int *colInd, const double *vals,
                  const double *b, double *y, double *x);
void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
