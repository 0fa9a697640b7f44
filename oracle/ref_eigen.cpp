// Linear solve of the oracle's Newton step through the reference's vendored
// Eigen 3.2.8 SparseLU with COLAMD ordering, the reference's Eigen_SparseLU
// option (src/powerflow/newtonpf.h:341-363): the matrix is inserted entry by
// entry from CSR, compressed, analysed and factorised on every call.
// Built only by oracle/Makefile's `ref` target, from the headers where they
// lie under /root/reference/src; the result goes to oracle/_ref/ and serves as
// a checker of the oracle's own sparse LU.  TEST INFRASTRUCTURE ONLY.
#include <Eigen/SparseLU>

extern "C" int orc_ref_linsolve(int n, const int *rowptr, const int *colind,
                                const double *val, const double *b,
                                double *x) {
  using namespace Eigen;
  SparseMatrix<double> A(n, n);
  for (int i = 0; i < n; i++)
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
      A.insert(i, colind[k]) = val[k];
  A.makeCompressed();
  SparseLU<SparseMatrix<double>, COLAMDOrdering<int> > lu;
  lu.compute(A);
  if (lu.info() != Success) return 2; /* singular */
  VectorXd B(n);
  for (int i = 0; i < n; i++) B(i) = b[i];
  VectorXd X = lu.solve(B);
  for (int i = 0; i < n; i++) x[i] = X(i);
  return 0;
}
