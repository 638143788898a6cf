// Substate text IO (include/cal2DIO.h), formats of
// OpenCAL-OMP/source/cal2DIO.c:22-76 and cal2DBufferIO.c:22-194: whitespace
// separated values, integers with "%d", reals with "%.9f", each followed by
// two spaces, one row per line.  Loading fills current and copies it into next.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine.hpp"

extern "C" {
#include "cal2DIO.h"
}

using namespace calhip;

template <typename T, typename Conv>
static void load_matrix(T* M, int rows, int columns, FILE* f, Conv conv)
{
	char tok[256];
	size_t n = (size_t)rows * columns;
	for (size_t c = 0; c < n; ++c) {
		if (fscanf(f, "%255s", tok) != 1) return;
		M[c] = (T)conv(tok);
	}
}

extern "C" void calfLoadMatrix2Db(CALbyte* M, int rows, int columns, FILE* f) { load_matrix(M, rows, columns, f, atoi); }
extern "C" void calfLoadMatrix2Di(CALint* M, int rows, int columns, FILE* f) { load_matrix(M, rows, columns, f, atoi); }
extern "C" void calfLoadMatrix2Dr(CALreal* M, int rows, int columns, FILE* f) { load_matrix(M, rows, columns, f, atof); }

template <typename T>
static void save_matrix(const T* M, int rows, int columns, FILE* f, const char* fmt)
{
	char s[64];
	for (int i = 0; i < rows; ++i) {
		for (int j = 0; j < columns; ++j) {
			snprintf(s, sizeof(s), fmt, M[(size_t)i * columns + j]);
			fprintf(f, "%s ", s);
		}
		fprintf(f, "\n");
	}
}

extern "C" void calfSaveMatrix2Db(CALbyte* M, int rows, int columns, FILE* f) { save_matrix(M, rows, columns, f, "%d "); }
extern "C" void calfSaveMatrix2Di(CALint* M, int rows, int columns, FILE* f) { save_matrix(M, rows, columns, f, "%d "); }
extern "C" void calfSaveMatrix2Dr(CALreal* M, int rows, int columns, FILE* f) { save_matrix(M, rows, columns, f, "%.9f "); }

#define PATH_FUNCS(SUF, T)                                                                  \
	extern "C" CALbyte calLoadMatrix2D##SUF(T* M, int rows, int columns, char* path)       \
	{                                                                                       \
		FILE* f = fopen(path, "r");                                                         \
		if (!f) return CAL_FALSE;                                                           \
		calfLoadMatrix2D##SUF(M, rows, columns, f);                                         \
		fclose(f);                                                                          \
		return CAL_TRUE;                                                                    \
	}                                                                                       \
	extern "C" CALbyte calSaveMatrix2D##SUF(T* M, int rows, int columns, char* path)       \
	{                                                                                       \
		FILE* f = fopen(path, "w");                                                         \
		if (!f) return CAL_FALSE;                                                           \
		calfSaveMatrix2D##SUF(M, rows, columns, f);                                         \
		fclose(f);                                                                          \
		return CAL_TRUE;                                                                    \
	}                                                                                       \
	extern "C" void calfLoadSubstate2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, FILE* f) \
	{                                                                                       \
		if (ensure_host_for(X(m), Q)) return;                                                      \
		calfLoadMatrix2D##SUF(Q->current, m->rows, m->columns, f);                          \
		if (Q->next) memcpy(Q->next, Q->current, sizeof(T) * (size_t)m->rows * m->columns); \
	}                                                                                       \
	extern "C" CALbyte calLoadSubstate2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, char* path) \
	{                                                                                       \
		if (ensure_host_for(X(m), Q)) return CAL_FALSE;                                            \
		CALbyte ok = calLoadMatrix2D##SUF(Q->current, m->rows, m->columns, path);           \
		if (Q->next) memcpy(Q->next, Q->current, sizeof(T) * (size_t)m->rows * m->columns); \
		return ok;                                                                          \
	}                                                                                       \
	extern "C" void calfSaveSubstate2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, FILE* f) \
	{                                                                                       \
		if (host_read(X(m), Q)) return;                                                     \
		calfSaveMatrix2D##SUF(Q->current, m->rows, m->columns, f);                          \
	}                                                                                       \
	extern "C" CALbyte calSaveSubstate2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, char* path) \
	{                                                                                       \
		if (host_read(X(m), Q)) return CAL_FALSE;                                           \
		return calSaveMatrix2D##SUF(Q->current, m->rows, m->columns, path);                 \
	}

PATH_FUNCS(b, CALbyte)
PATH_FUNCS(i, CALint)
PATH_FUNCS(r, CALreal)
