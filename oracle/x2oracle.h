/* CPU restatement of the x2dhf relaxation kernels (test infrastructure only;
 * see x2oracle.c). */
#ifndef X2ORACLE_H
#define X2ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
void x2o_putin(int nni, int nmu, int isym, const double *fun, double *work);
void x2o_putout(int nni, int nmu, double *fun, const double *work);
void x2o_boundary(int nni, int nmu, int isym, double *f, const double *exeven);
void x2o_mcsor(int nni, int nmu, int isym, double *work, const double *lhs, const double *rhs,
               const double *b, const double *d, const double *coef, const double *exeven,
               double omega, int nsweeps, int mode);
void x2o_mcsor_f32(int nni, int nmu, int isym, float *f, const float *lhs, const float *rhs,
                   const float *b, const float *d, const double *coef, const double *exeven,
                   double omega, int nsweeps);
void x2o_residual(int nni, int nmu, const double *work, const double *lhs, const double *rhs,
                  const double *b, const double *d, const double *coef, double *res);
void x2o_orb_asympt_get(int nni, int nmu, const double *vxisq, const double *vetasq,
                        const double *fa, double *edecay);
void x2o_orb_asympt_set(int nni, int nmu, double *psi, const double *edecay);
long x2o_mesh(int nni, int nmu, int ordering, int *seq);
void x2o_relax_ordered(int nni, int nmu, int isym, double *work, const double *lhs, const double *rhs,
                       const double *b, const double *d, const double *coef, const double *exeven,
                       double omega, int nsweeps, int mode, int ordering, int *seq);
#ifdef __cplusplus
}
#endif
#endif
