/* Slice-major 3D host buffers (OpenCAL-OMP/include/cal3DBuffer.h:20-136).
 * Element (i, j, k) of a rows x columns x slices buffer is
 * M[k*rows*columns + i*columns + j]. */
#ifndef cal3DBuffer_h
#define cal3DBuffer_h

#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cal3DBuffer.c:19-40 */
CALbyte* calAllocBuffer3Db(int rows, int columns, int slices);
CALint* calAllocBuffer3Di(int rows, int columns, int slices);
CALreal* calAllocBuffer3Dr(int rows, int columns, int slices);
void calDeleteBuffer3Db(CALbyte* M);
void calDeleteBuffer3Di(CALint* M);
void calDeleteBuffer3Dr(CALreal* M);
/* cal3DBuffer.c:43-163 */
void calCopyBuffer3Db(CALbyte* M_src, CALbyte* M_dest, int rows, int columns, int slices);
void calCopyBuffer3Di(CALint* M_src, CALint* M_dest, int rows, int columns, int slices);
void calCopyBuffer3Dr(CALreal* M_src, CALreal* M_dest, int rows, int columns, int slices);
void calCopyActiveCellsBuffer3Db(CALbyte* M_src, CALbyte* M_dest, int rows, int columns, int slices,
                                 struct CALCell3D* active_cells, int sizeof_active_cells);
void calCopyActiveCellsBuffer3Di(CALint* M_src, CALint* M_dest, int rows, int columns, int slices,
                                 struct CALCell3D* active_cells, int sizeof_active_cells);
void calCopyActiveCellsBuffer3Dr(CALreal* M_src, CALreal* M_dest, int rows, int columns, int slices,
                                 struct CALCell3D* active_cells, int sizeof_active_cells);
/* cal3DBuffer.c:166-210 */
void calAddBuffer3Db(CALbyte* M_op1, CALbyte* M_op2, CALbyte* M_dest, int rows, int columns, int slices);
void calAddBuffer3Di(CALint* M_op1, CALint* M_op2, CALint* M_dest, int rows, int columns, int slices);
void calAddBuffer3Dr(CALreal* M_op1, CALreal* M_op2, CALreal* M_dest, int rows, int columns, int slices);
void calSubtractBuffer3Db(CALbyte* M_op1, CALbyte* M_op2, CALbyte* M_dest, int rows, int columns, int slices);
void calSubtractBuffer3Di(CALint* M_op1, CALint* M_op2, CALint* M_dest, int rows, int columns, int slices);
void calSubtractBuffer3Dr(CALreal* M_op1, CALreal* M_op2, CALreal* M_dest, int rows, int columns, int slices);
/* cal3DBuffer.c:214-250 (integer fills are memset, as in the reference) */
void calSetBuffer3Db(CALbyte* M, int rows, int columns, int slices, CALbyte value);
void calSetBuffer3Di(CALint* M, int rows, int columns, int slices, CALint value);
void calSetBuffer3Dr(CALreal* M, int rows, int columns, int slices, CALreal value);
void calSetActiveCellsBuffer3Db(CALbyte* M, int rows, int columns, int slices, CALbyte value,
                                struct CALCell3D* active_cells, int sizeof_active_cells);
void calSetActiveCellsBuffer3Di(CALint* M, int rows, int columns, int slices, CALint value,
                                struct CALCell3D* active_cells, int sizeof_active_cells);
void calSetActiveCellsBuffer3Dr(CALreal* M, int rows, int columns, int slices, CALreal value,
                                struct CALCell3D* active_cells, int sizeof_active_cells);

#define calSetBuffer3DElement(M, rows, columns, i, j, k, value) \
	((M)[(((k) * (rows) * (columns)) + ((i) * (columns)) + (j))] = (value))
#define calGetBuffer3DElement(M, rows, columns, i, j, k) (M[(((k) * (rows) * (columns)) + ((i) * (columns)) + (j))])

#ifdef __cplusplus
}
#endif
#endif
