/*
 * rg_oracle.h -- shared declarations of the FP64 CPU oracle (test infrastructure only; see
 * rg_oracle.c).
 */
#ifndef RG_ORACLE_H
#define RG_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double complex zc;

#define RGO_PI 3.14159265358979323846264338327950288

enum { RGO_OK = 0, RGO_EINVAL = 1, RGO_ESINGULAR = 2, RGO_ENOCONV = 3, RGO_ENOMEM = 4 };
enum { RGO_DIRICHLET = 0, RGO_PERIODIC = 1 };

typedef struct {
    int boundtype;
    int nxdiso, nxfree, nx;
    double wol, exfac, aratio, dscat, dabso, napera, naperb;
    int nmu, ny, imin, imax, jmin, jmax, neigv, nprop; /* imin.. are 1-based like the reference */
    double kl, kdx, dos, stdu;
    int proj_method; /* 0: direct modal^H diag(u) modal (the reference's form); 1: transform form */
    double *kpmesh, *ymesh;
    zc *mumesh, *bmat, *abound, *abulk, *modal; /* modal: ny x nmu */
} rgo_wg;


void rgo_slice_matrix(const rgo_wg *w, int ix, const double *upot, zc *A, zc *work);
int rgo_inverse(int n, zc *A, int *ipiv, zc *work);
int rgo_singular_values(int m, int n, const zc *A, double *S);

#endif
