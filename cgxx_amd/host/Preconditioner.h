// Jacobi preconditioner (interface of the reference's Preconditioner.h:12-25).
#ifndef CGHIP_HOST_PRECONDITIONER_H
#define CGHIP_HOST_PRECONDITIONER_H

#include "Matrix.h"
#include "def.h"

struct Jacobi {
  // C[i] = 1 / A_ii
  floatType *C = nullptr;

  virtual ~Jacobi() {}
  // Fills C from the diagonal entries of coo; a row without a diagonal entry
  // is an error (the reference leaves C[i] uninitialised there).
  void init(const MatrixCOO &coo);
  virtual void allocateC(int N);
  virtual void deallocateC();
};

#endif
