/* Substate text IO (OpenCAL-OMP/include/cal2DIO.h:21-71); the matrix
 * functions of cal2DBufferIO.h come with it.
 * Whitespace-separated matrices; reals are written with "%.9f" as the
 * OpenMP reference does (OpenCAL-OMP/source/cal2DBufferIO.c:142).  The save
 * functions first bring the substate back from the GPU when a device run
 * left it there. */
#ifndef cal2DIO_h
#define cal2DIO_h

#include <stdio.h>
#include <cal2D.h>
#include <cal2DBufferIO.h>

#ifdef __cplusplus
extern "C" {
#endif

void calfLoadSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, FILE* f);
void calfLoadSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, FILE* f);
void calfLoadSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, FILE* f);
CALbyte calLoadSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, char* path);
CALbyte calLoadSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, char* path);
CALbyte calLoadSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, char* path);
void calfSaveSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, FILE* f);
void calfSaveSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, FILE* f);
void calfSaveSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, FILE* f);
CALbyte calSaveSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, char* path);
CALbyte calSaveSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, char* path);
CALbyte calSaveSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, char* path);

#ifdef __cplusplus
}
#endif
#endif
