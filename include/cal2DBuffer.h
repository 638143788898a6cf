/* Row-major 2D host buffers (OpenCAL-OMP/include/cal2DBuffer.h:20-136).
 * Element (i, j) of a rows x columns matrix is M[i*columns + j]. */
#ifndef cal2DBuffer_h
#define cal2DBuffer_h

#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

CALbyte* calAllocBuffer2Db(int rows, int columns);
CALint* calAllocBuffer2Di(int rows, int columns);
CALreal* calAllocBuffer2Dr(int rows, int columns);
void calDeleteBuffer2Db(CALbyte* M);
void calDeleteBuffer2Di(CALint* M);
void calDeleteBuffer2Dr(CALreal* M);
void calCopyBuffer2Db(CALbyte* M_src, CALbyte* M_dest, int rows, int columns);
void calCopyBuffer2Di(CALint* M_src, CALint* M_dest, int rows, int columns);
void calCopyBuffer2Dr(CALreal* M_src, CALreal* M_dest, int rows, int columns);
void calCopyActiveCellsBuffer2Db(CALbyte* M_src, CALbyte* M_dest, int rows, int columns,
                                 struct CALCell2D* active_cells, int sizeof_active_cells);
void calCopyActiveCellsBuffer2Di(CALint* M_src, CALint* M_dest, int rows, int columns,
                                 struct CALCell2D* active_cells, int sizeof_active_cells);
void calCopyActiveCellsBuffer2Dr(CALreal* M_src, CALreal* M_dest, int rows, int columns,
                                 struct CALCell2D* active_cells, int sizeof_active_cells);
void calAddBuffer2Db(CALbyte* M_op1, CALbyte* M_op2, CALbyte* M_dest, int rows, int columns);
void calAddBuffer2Di(CALint* M_op1, CALint* M_op2, CALint* M_dest, int rows, int columns);
void calAddBuffer2Dr(CALreal* M_op1, CALreal* M_op2, CALreal* M_dest, int rows, int columns);
void calSubtractBuffer2Db(CALbyte* M_op1, CALbyte* M_op2, CALbyte* M_dest, int rows, int columns);
void calSubtractBuffer2Di(CALint* M_op1, CALint* M_op2, CALint* M_dest, int rows, int columns);
void calSubtractBuffer2Dr(CALreal* M_op1, CALreal* M_op2, CALreal* M_dest, int rows, int columns);
void calSetBuffer2Db(CALbyte* M, int rows, int columns, CALbyte value);
void calSetBuffer2Di(CALint* M, int rows, int columns, CALint value);
void calSetBuffer2Dr(CALreal* M, int rows, int columns, CALreal value);
void calSetActiveCellsBuffer2Db(CALbyte* M, int rows, int columns, CALbyte value,
                                struct CALCell2D* active_cells, int sizeof_active_cells);
void calSetActiveCellsBuffer2Di(CALint* M, int rows, int columns, CALint value,
                                struct CALCell2D* active_cells, int sizeof_active_cells);
void calSetActiveCellsBuffer2Dr(CALreal* M, int rows, int columns, CALreal value,
                                struct CALCell2D* active_cells, int sizeof_active_cells);

#define calSetMatrixElement(M, columns, i, j, value) ((M)[(((i) * (columns)) + (j))] = (value))
#define calGetMatrixElement(M, columns, i, j) (M[(((i) * (columns)) + (j))])

#ifdef __cplusplus
}
#endif
#endif
