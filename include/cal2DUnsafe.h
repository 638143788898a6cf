/* Unsafe operations of OpenCAL's 2D API (OpenCAL-OMP/include/cal2DUnsafe.h,
 * OpenCAL-OMP/source/cal2DUnsafe.c:17-275): a process writes a neighbour's
 * cell, reads the next layer, or adds to the next layer.
 *
 * On the host they act on the host mirrors with the reference semantics.
 * In process source run by the generic device program (calhip2D.h) they are
 * device functions: calAddNext* adds atomically (the reference takes a cell
 * lock), so additions of real values from several cells land in an
 * unspecified order, as with the reference's OpenMP threads. */
#ifndef cal2DUnsafe_h
#define cal2DUnsafe_h

#include <cal2D.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cal2DUnsafe.c:17-68 — current and next of neighbour n. */
void calInitX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n, CALbyte value);
void calInitX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n, CALint value);
void calInitX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n, CALreal value);

/* cal2DUnsafe.c:70-105 — the next layer of (i, j). */
CALbyte calGetNext2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j);
CALint calGetNext2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j);
CALreal calGetNext2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j);

/* cal2DUnsafe.c:107-140 — the next layer of neighbour n. */
CALbyte calGetNextX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n);
CALint calGetNextX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n);
CALreal calGetNextX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n);

/* cal2DUnsafe.c:142-175 — next of neighbour n. */
void calSetX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n, CALbyte value);
void calSetX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n, CALint value);
void calSetX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n, CALreal value);

/* cal2DUnsafe.c:176-211 — current of neighbour n. */
void calSetCurrentX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n, CALbyte value);
void calSetCurrentX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n, CALint value);
void calSetCurrentX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n, CALreal value);

/* cal2DUnsafe.c:213-275 — next += value at (i, j) or at neighbour n. */
void calAddNext2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, CALbyte value);
void calAddNext2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, CALint value);
void calAddNext2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, CALreal value);
void calAddNextX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n, CALbyte value);
void calAddNextX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n, CALint value);
void calAddNextX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n, CALreal value);

#ifdef __cplusplus
}
#endif

#endif
