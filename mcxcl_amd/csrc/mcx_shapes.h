/*
 * mcx_shapes.h - JSON shape rasteriser (reference src/mcx_shapes.h).
 */
#ifndef MCX_SHAPES_H
#define MCX_SHAPES_H

#include "mcx_utils.h"

#ifdef __cplusplus
extern "C" {
#endif

/* shapesnode: an MJNode* holding either a {"Shapes":[...]} object, the Shapes
   array itself, or an uncompressed JData volume object.  Returns 0 on success,
   non-zero with a message in err (at least 256 bytes) otherwise. */
int mcx_rasterize_shapes(const void* shapesnode, unsigned int** vol, MCXUInt4* dim, int issrcfrom0, int rowmajor,
                         char* err);

#ifdef __cplusplus
}
#endif

#endif
