// Substate and buffer text IO of the 3D API (include/cal3DIO.h,
// cal3DBufferIO.h), formats of OpenCAL-OMP/source/cal3DBufferIO.c:22-206
// and cal3DIO.c:17-76: whitespace-separated values read slice by slice;
// written as "%d " / "%f " each followed by a space, a newline after each
// row and an empty line after each slice.  Loading fills current and copies
// it into next.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine.hpp"

extern "C" {
#include "cal3DIO.h"
}

using namespace calhip;

template <typename T, typename Conv>
static void load3(T* M, int rows, int columns, int slices, FILE* f, Conv conv)
{
	char tok[256];
	const size_t n = (size_t)rows * columns * slices;
	for (size_t c = 0; c < n; ++c) {
		if (fscanf(f, "%255s", tok) != 1) return;
		M[c] = (T)conv(tok);
	}
}

template <typename T>
static void save3(const T* M, int rows, int columns, int slices, FILE* f, const char* fmt)
{
	char s[64];
	for (int k = 0; k < slices; ++k) {
		for (int i = 0; i < rows; ++i) {
			for (int j = 0; j < columns; ++j) {
				snprintf(s, sizeof(s), fmt, M[((size_t)k * rows + i) * columns + j]);
				fprintf(f, "%s ", s);
			}
			fprintf(f, "\n");
		}
		fprintf(f, "\n");
	}
}

extern "C" void calfLoadBuffer3Db(CALbyte* M, int rows, int columns, int slices, FILE* f) { load3(M, rows, columns, slices, f, atoi); }
extern "C" void calfLoadBuffer3Di(CALint* M, int rows, int columns, int slices, FILE* f) { load3(M, rows, columns, slices, f, atoi); }
extern "C" void calfLoadBuffer3Dr(CALreal* M, int rows, int columns, int slices, FILE* f) { load3(M, rows, columns, slices, f, atof); }
extern "C" void calfSaveBuffer3Db(CALbyte* M, int rows, int columns, int slices, FILE* f) { save3(M, rows, columns, slices, f, "%d "); }
extern "C" void calfSaveBuffer3Di(CALint* M, int rows, int columns, int slices, FILE* f) { save3(M, rows, columns, slices, f, "%d "); }
extern "C" void calfSaveBuffer3Dr(CALreal* M, int rows, int columns, int slices, FILE* f) { save3(M, rows, columns, slices, f, "%f "); }

#define IO3_FUNCS(SUF, T)                                                                          \
	extern "C" CALbyte calLoadBuffer3D##SUF(T* M, int rows, int columns, int slices, char* path)    \
	{                                                                                              \
		FILE* f = fopen(path, "r");                                                                \
		if (!f) return CAL_FALSE;                                                                  \
		calfLoadBuffer3D##SUF(M, rows, columns, slices, f);                                        \
		fclose(f);                                                                                 \
		return CAL_TRUE;                                                                           \
	}                                                                                              \
	extern "C" CALbyte calSaveBuffer3D##SUF(T* M, int rows, int columns, int slices, char* path)    \
	{                                                                                              \
		FILE* f = fopen(path, "w");                                                                \
		if (!f) return CAL_FALSE;                                                                  \
		calfSaveBuffer3D##SUF(M, rows, columns, slices, f);                                        \
		fclose(f);                                                                                 \
		return CAL_TRUE;                                                                           \
	}                                                                                              \
	extern "C" void calfLoadSubstate3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, FILE* f) \
	{                                                                                              \
		if (ensure_host_for(&X3(m)->core, Q)) return;                                              \
		calfLoadBuffer3D##SUF(Q->current, m->rows, m->columns, m->slices, f);                      \
		if (Q->next) memcpy(Q->next, Q->current, sizeof(T) * (size_t)m->rows * m->columns * m->slices); \
	}                                                                                              \
	extern "C" CALbyte calLoadSubstate3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, char* path) \
	{                                                                                              \
		if (ensure_host_for(&X3(m)->core, Q)) return CAL_FALSE;                                    \
		CALbyte ok = calLoadBuffer3D##SUF(Q->current, m->rows, m->columns, m->slices, path);       \
		if (Q->next) memcpy(Q->next, Q->current, sizeof(T) * (size_t)m->rows * m->columns * m->slices); \
		return ok;                                                                                 \
	}                                                                                              \
	extern "C" void calfSaveSubstate3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, FILE* f) \
	{                                                                                              \
		if (host_read(&X3(m)->core, Q)) return;                                                    \
		calfSaveBuffer3D##SUF(Q->current, m->rows, m->columns, m->slices, f);                      \
	}                                                                                              \
	extern "C" CALbyte calSaveSubstate3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, char* path) \
	{                                                                                              \
		if (host_read(&X3(m)->core, Q)) return CAL_FALSE;                                          \
		return calSaveBuffer3D##SUF(Q->current, m->rows, m->columns, m->slices, path);             \
	}

IO3_FUNCS(b, CALbyte)
IO3_FUNCS(i, CALint)
IO3_FUNCS(r, CALreal)
