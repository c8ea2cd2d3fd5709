/* Minimal declarations of the PETSc and reference types and functions the
 * reference-side binding in INTEGRATION.md uses, so that the binding can be
 * compiled (not linked or run) without PETSc: the types carry only the
 * fields the binding touches, with the reference's names and C types
 * (variables.h; PetscInt is a 32-bit int, PetscReal/PetscScalar double). */
#ifndef PETSC_STUB_H
#define PETSC_STUB_H

typedef int PetscErrorCode;
typedef int PetscInt;
typedef double PetscReal;
typedef double PetscScalar;
typedef int PetscBool;
typedef struct _p_Vec *Vec;
typedef struct _p_DM *DM;
typedef struct { PetscInt mx, my, mz; } DMDALocalInfo;
typedef struct { PetscReal x, y, z; } Cmpnts;

typedef struct {
  DM da;
  DMDALocalInfo info;
  PetscInt bctype[6];
  PetscReal ren, Ma, Pr, gamma, dt, st;
  PetscInt ti;
  Vec Q, P, lQ, Nvert;
} UserCtx;

typedef struct {
  PetscInt n_v, n_elmt;
  PetscReal *x_bp, *y_bp, *z_bp;
  PetscInt *nv1, *nv2, *nv3;
  Cmpnts *u, *uold;
} IBMNodes;

typedef struct {
  PetscReal x_c, y_c, z_c;
  PetscReal F_x, F_y, F_z, A_tot, M_x, M_y, M_z, Mdpdn_x, Mdpdn_y, Mdpdn_z;
} FSInfo;

/* globals of main.c the binding reads */
extern PetscInt les, invicid, central, isothermal, tistart, rstart_flg;
extern PetscInt i_homo_filter, j_homo_filter, k_homo_filter;

PetscErrorCode PetscOptionsGetInt(void *options, const char *pre, const char *name, PetscInt *v, PetscBool *set);
PetscErrorCode PetscOptionsGetReal(void *options, const char *pre, const char *name, PetscReal *v, PetscBool *set);
PetscErrorCode PetscOptionsGetBool(void *options, const char *pre, const char *name, PetscBool *v, PetscBool *set);
PetscErrorCode PetscPrintf(int comm, const char *format, ...);
PetscErrorCode DMGetCoordinates(DM dm, Vec *c);
PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecGetArray(Vec v, PetscScalar **a);
PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a);
PetscErrorCode PetscError(int comm, int line, const char *func, const char *file, int n, int p, const char *mess, ...);
#define PETSC_COMM_SELF 0
#define PETSC_COMM_WORLD 0
#define SETERRQ1(comm, n, s, a1) return PetscError(comm, __LINE__, __func__, __FILE__, n, 1, s, a1)

#endif
