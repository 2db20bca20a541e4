/*
 * petsc.h -- test double of the few PETSc declarations that
 * include/amat_amd/aMatFree.hpp uses under BUILD_WITH_PETSC, with just enough
 * behaviour to replay the calls a PETSc solver makes on the shell matrix
 * (MatCreateShell / MatShellSetOperation / MatShellGetContext, MatMult,
 * MatGetDiagonal, MatGetDiagonalBlock, and arrays of vectors).  It is NOT
 * PETSc: PETSc is not installed here.  Only the programs under tests/cpp
 * include it; MPI comes from the real mpi.h.
 */
#ifndef AMAT_AMD_TEST_PETSC_STUB_H
#define AMAT_AMD_TEST_PETSC_STUB_H

#include <mpi.h>

#include <vector>

typedef int PetscErrorCode;
typedef int PetscInt;
typedef double PetscScalar;
typedef double PetscReal;

#define PETSC_DETERMINE (-1)
#define PETSC_ERR_SUP 56
#define PETSC_ERR_LIB 76
#define PETSC_COMM_SELF MPI_COMM_SELF
#define PETSC_COMM_WORLD MPI_COMM_WORLD

typedef enum { MATOP_MULT = 0, MATOP_GET_DIAGONAL = 1, MATOP_GET_DIAGONAL_BLOCK = 2 } MatOperation;

struct _p_Mat {
    MPI_Comm comm;
    PetscInt m, n;
    void* ctx;
    void (*ops[3])(void);
};
typedef struct _p_Mat* Mat;

struct _p_Vec {
    std::vector<PetscScalar> a;
    int borrowed;  // arrays handed out and not yet restored
};
typedef struct _p_Vec* Vec;

inline PetscErrorCode MatCreateShell(MPI_Comm comm, PetscInt m, PetscInt n, PetscInt, PetscInt, void* ctx, Mat* A)
{
    *A = new _p_Mat{comm, m, n, ctx, {nullptr, nullptr, nullptr}};
    return 0;
}
inline PetscErrorCode MatShellSetOperation(Mat A, MatOperation op, void (*f)(void))
{
    A->ops[op] = f;
    return 0;
}
/* the current PETSc form: ctx is the address of the caller's pointer */
inline PetscErrorCode MatShellGetContext(Mat A, void* ctx)
{
    *static_cast<void**>(ctx) = A->ctx;
    return 0;
}
inline PetscErrorCode MatDestroy(Mat* A)
{
    delete *A;
    *A = nullptr;
    return 0;
}
inline PetscErrorCode MatMult(Mat A, Vec x, Vec y)
{
    if (!A->ops[MATOP_MULT]) return PETSC_ERR_SUP;
    return reinterpret_cast<PetscErrorCode (*)(Mat, Vec, Vec)>(A->ops[MATOP_MULT])(A, x, y);
}
inline PetscErrorCode MatGetDiagonal(Mat A, Vec d)
{
    if (!A->ops[MATOP_GET_DIAGONAL]) return PETSC_ERR_SUP;
    return reinterpret_cast<PetscErrorCode (*)(Mat, Vec)>(A->ops[MATOP_GET_DIAGONAL])(A, d);
}
inline PetscErrorCode MatGetDiagonalBlock(Mat A, Mat* a)
{
    if (!A->ops[MATOP_GET_DIAGONAL_BLOCK]) return PETSC_ERR_SUP;
    return reinterpret_cast<PetscErrorCode (*)(Mat, Mat*)>(A->ops[MATOP_GET_DIAGONAL_BLOCK])(A, a);
}

inline PetscErrorCode VecCreateMPI(MPI_Comm, PetscInt n, PetscInt, Vec* v)
{
    *v = new _p_Vec{std::vector<PetscScalar>(n, 0.0), 0};
    return 0;
}
inline PetscErrorCode VecDestroy(Vec* v)
{
    delete *v;
    *v = nullptr;
    return 0;
}
inline PetscErrorCode VecGetArray(Vec v, PetscScalar** a)
{
    v->borrowed++;
    *a = v->a.data();
    return 0;
}
inline PetscErrorCode VecRestoreArray(Vec v, PetscScalar** a)
{
    v->borrowed--;
    *a = nullptr;
    return 0;
}
inline PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar** a)
{
    v->borrowed++;
    *a = v->a.data();
    return 0;
}
inline PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar** a)
{
    v->borrowed--;
    *a = nullptr;
    return 0;
}

#endif /* AMAT_AMD_TEST_PETSC_STUB_H */
