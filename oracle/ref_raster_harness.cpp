// ref_raster_harness.cpp -- test-only C entry point around the reference's
// own PuresoftRasterizer (src/puresoft3d/rasterizer.{h,cpp}), compiled from
// the reference tree by oracle/Makefile into oracle/_ref/libref_raster.so.
// It lets the tests compare the oracle's rasterizer restatement with the
// reference bit for bit.  Nothing here ships in the product library.
#include <cstring>
#include <climits>
#include "rasterizer.h"

extern "C" int ref_push_triangle(int width, int height, const float* v0, const float* v1,
                                 const float* v2, int* first_row, int* last_row, int* rows_out,
                                 float* verts_out)
{
    // One rasterizer per call: PuresoftRasterizer::initialize (rasterizer.cpp:75-93)
    PuresoftRasterizer rast;
    const PuresoftRasterizer::RESULT* res = rast.initialize(width, height);
    // Fill every row with a sentinel so rows this triangle does not write stay visible.
    PuresoftRasterizer::RESULT_ROW* rows = res->m_rows;
    for (int i = 0; i < height; i++) {
        rows[i].left = rows[i].leftClamped = rows[i].right = rows[i].rightClamped = INT_MAX;
        rows[i].leftVerts[0] = rows[i].leftVerts[1] = rows[i].rightVerts[0] = rows[i].rightVerts[1] = -7;
    }
    bool ok = rast.pushTriangle(v0, v1, v2);
    *first_row = res->firstRow;
    *last_row = res->lastRow;
    for (int i = 0; i < 3; i++) {
        verts_out[2 * i] = res->vertices[i].x;
        verts_out[2 * i + 1] = res->vertices[i].y;
    }
    for (int i = 0; i < height; i++) {
        int* o = rows_out + 8 * i;
        o[0] = rows[i].left; o[1] = rows[i].leftClamped;
        o[2] = rows[i].leftVerts[0]; o[3] = rows[i].leftVerts[1];
        o[4] = rows[i].right; o[5] = rows[i].rightClamped;
        o[6] = rows[i].rightVerts[0]; o[7] = rows[i].rightVerts[1];
    }
    return ok ? 1 : 0;
}
