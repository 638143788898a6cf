/* Unsafe operations of OpenCAL's 3D API (OpenCAL-OMP/include/cal3DUnsafe.h,
 * OpenCAL-OMP/source/cal3DUnsafe.c:18-244): a process writes a neighbour's
 * cell, reads the next layer, or adds to the next layer.
 *
 * On the host they act on the host mirrors with the reference semantics.
 * In process source run by the generic device program (calhip3D.h) they are
 * device functions: calAddNext* adds atomically (the reference takes a cell
 * lock), so additions of real values from several cells land in an
 * unspecified order, as with the reference's OpenMP threads. */
#ifndef cal3DUnsafe_h
#define cal3DUnsafe_h

#include <cal3D.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cal3DUnsafe.c:18-59 — current and next of neighbour n. */
void calInitX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n, CALbyte value);
void calInitX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n, CALint value);
void calInitX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n, CALreal value);

/* cal3DUnsafe.c:62-97 — the next layer of (i, j, k). */
CALbyte calGetNext3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k);
CALint calGetNext3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k);
CALreal calGetNext3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k);

/* cal3DUnsafe.c:100-129 — the next layer of neighbour n. */
CALbyte calGetNextX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n);
CALint calGetNextX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n);
CALreal calGetNextX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n);

/* cal3DUnsafe.c:132-161 — write the next layer of neighbour n. */
void calSetX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n, CALbyte value);
void calSetX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n, CALint value);
void calSetX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n, CALreal value);

/* cal3DUnsafe.c:164-186 — write the current layer of neighbour n. */
void calSetCurrentX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n,
                       CALbyte value);
void calSetCurrentX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n, CALint value);
void calSetCurrentX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n,
                       CALreal value);

/* cal3DUnsafe.c:188-244 — add to the next layer of (i, j, k) / of neighbour
 * n (defined there though its header does not declare them). */
void calAddNext3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, CALbyte value);
void calAddNext3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, CALint value);
void calAddNext3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, CALreal value);
void calAddNextX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n, CALbyte value);
void calAddNextX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n, CALint value);
void calAddNextX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n, CALreal value);

#ifdef __cplusplus
}
#endif
#endif
