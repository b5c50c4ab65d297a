/* raishin_io.h -- the reference's restart and analysis file formats.
 *
 * restar.f90 writes Fortran unformatted sequential files: every WRITE is one
 * record, framed by its byte length as a 4-byte little-endian integer before
 * and after the payload (the flang/gfortran default).  Integers are 4 bytes,
 * reals 8.  These entry points write and read the same bytes from C, so a
 * driver that keeps its state on the device (raishin_model, batches,
 * serving) checkpoints and restarts in the reference's own files, and a
 * Fortran build of the reference reads them (icres=1, main.f90:560-571).
 *
 * The first three take host arrays in the reference layout with the
 * reference's argument lists (integers by address, as Fortran passes them),
 * plus the file path in front: NULL means the reference's own file name
 * (restartRRR.outdat, structrRRRR-NNNN.outdat).  They are void and report
 * through the sticky error (raishin_last_error).  They run on the host only.
 *
 * One rank: output1's index range is that of rank 0 of a 1x1x1 grid
 * (myranki = myrankj = myrankk = 0; other values are rejected).
 */
#ifndef RAISHIN_IO_H
#define RAISHIN_IO_H
#include "raishin_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

/* store1 (restar.f90:5-88): records (myrank, nd), (it0, ih0), (time), then
 * for k, j, i (i fastest) four records of nv=9 reals: uu, u0, uri, uri0. */
void raishin_store1(const char *path, const double *uu, const double *u0, const double *uri, const double *uri0,
                    const int *it0, const int *ih0, const double *time, const int *nd, const int *npe,
                    const int *myrank, const int *is1, const int *ie1, const int *js1, const int *je1,
                    const int *ks1, const int *ke1);

/* restar2 (restar.f90:273-346): reads what store1 wrote.  Every record
 * length is checked; a file written by another rank (restar.f90:323-326),
 * of another size, or cut short sets the error and leaves the arrays as far
 * as they were read. */
void raishin_restar2(const char *path, double *uu, double *u0, double *uri, double *uri0, int *it0, int *ih0,
                     double *time, int *nd, const int *npe, const int *myrank, const int *is1, const int *ie1,
                     const int *js1, const int *je1, const int *ks1, const int *ke1);

/* output1 (restar.f90:91-185): records (time, myrank, npe), (is, ie, js, je,
 * ks, ke), then for i, j, k (k fastest, restar.f90:171-179) one record of 13
 * reals: x1(i), x2(j), x3(k), uri(1:9,i,j,k), detg(i,j,k). */
void raishin_output1(const char *path, const double *uri, const double *detg, const double *x1, const double *x2,
                     const double *x3, const double *time, const int *nm1, const int *nd, const int *npe,
                     const int *myrank, const int *myranki, const int *myrankj, const int *myrankk,
                     const int *is1, const int *ie1, const int *js1, const int *je1, const int *ks1,
                     const int *ke1);

/* ---- the same files from a device-resident model ----------------------- */
/* The model's uri, uu and time are pulled from the device.  u0 and uri0 are
 * main.f90's copies of the initial state, which the time step never reads;
 * NULL writes the current state in their place.  Return 0, or -1 with the
 * error set. */
int raishin_model_store1(raishin_model *m, const char *path, const double *u0, const double *uri0, int it0,
                         int ih0, int nd, int myrank);
/* Loads uri, uu and time into the model (raishin_model_set_state); u0/uri0
 * (each 9*imax*jmax*kmax, may be NULL), it0, ih0 and nd (may be NULL) receive
 * the rest.  On error the model is unchanged. */
int raishin_model_restar2(raishin_model *m, const char *path, double *u0, double *uri0, int *it0, int *ih0,
                          int *nd, int myrank);
/* The analysis file of the model's current state; detg(imax,jmax,kmax) and the
 * centre positions x1, x2, x3 are the host arrays of the set-up. */
int raishin_model_output1(raishin_model *m, const char *path, const double *detg, const double *x1,
                          const double *x2, const double *x3, int nd, int myrank, int npe);

/* The model's pram (for drivers that size host buffers). */
void raishin_model_get_pram(raishin_model *m, raishin_pram *p);

#ifdef __cplusplus
}
#endif
#endif
