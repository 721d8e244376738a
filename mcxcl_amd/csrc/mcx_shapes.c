/*
 * mcx_shapes.c - rasterise JSON "Shapes" objects into the label volume.
 *
 * Semantics follow the reference rasteriser (src/mcx_shapes.c:55-860):
 *  - the volume starts as all label 1 when the domain size is already known
 *    (src/mcx_shapes.c:164-170); "Grid" re-creates it with a new size and tag;
 *  - "Origin" shifts the coordinates of Subgrid/Box/Cylinder (default origin is
 *    (1,1,1) for 1-based inputs and (0,0,0) when OriginType/issrcfrom0 is set);
 *  - "Sphere" tests voxel centres (i+0.5, j+0.5, k+0.5) against |c-O| <= R
 *    without the origin shift (src/mcx_shapes.c:249-263);
 *  - "Box" and "UpperSpace" test voxel centres, "Subgrid" uses 1-based integer
 *    corners, "Cylinder" tests the projection on the C0->C1 axis;
 *  - "XLayers/YLayers/ZLayers" take [start,end,tag] triplets (1-based,
 *    inclusive); "XSlabs/YSlabs/ZSlabs" take a "Bound" pair or list of pairs and
 *    a "Tag".  The reference's slab reader mis-reads the Bound member
 *    (src/mcx_shapes.c:442-470 walks the object's first member instead of the
 *    Bound array); this implementation follows the documented intent.
 * Only the first member of each element of the Shapes array is read, as in
 * the reference (src/mcx_shapes.c:177-199).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mcx_json.h"
#include "mcx_shapes.h"

typedef struct {
    unsigned int** vol;
    MCXUInt4* dim;
    float orig[3];
    int rowmajor;
} ShapeGrid;

static size_t vox_index(const ShapeGrid* g, int i, int j, int k) {
    if (g->rowmajor) {
        return (size_t)i * g->dim->y * g->dim->z + (size_t)j * g->dim->z + k;
    }

    return (size_t)k * g->dim->x * g->dim->y + (size_t)j * g->dim->x + i;
}

static int get_triplet(const MJNode* v, float out[3]) {
    if (!v || v->type != MJ_ARRAY || v->count != 3) {
        return 1;
    }

    for (int i = 0; i < 3; i++) {
        out[i] = (float)mj_num(mj_at(v, i), 0.0);
    }

    return 0;
}

static int raster_origin(const MJNode* obj, ShapeGrid* g, char* err) {
    if (get_triplet(obj, g->orig)) {
        snprintf(err, 256, "An Origin record does not contain a triplet");
        return 1;
    }

    return 0;
}

static int raster_grid(const MJNode* obj, ShapeGrid* g, char* err) {
    float sz[3];

    if (get_triplet(mj_get(obj, "Size"), sz)) {
        snprintf(err, 256, "A Grid command misses Size field");
        return 1;
    }

    g->dim->x = (unsigned int)sz[0];
    g->dim->y = (unsigned int)sz[1];
    g->dim->z = (unsigned int)sz[2];
    size_t n = (size_t)g->dim->x * g->dim->y * g->dim->z;
    free(*g->vol);
    *g->vol = n ? (unsigned int*)calloc(n, sizeof(unsigned int)) : NULL;
    const MJNode* tag = mj_get(obj, "Tag");

    if (tag && *g->vol) {
        unsigned int t = (unsigned int)mj_int(tag, 0);

        for (size_t i = 0; i < n; i++) {
            (*g->vol)[i] = t;
        }
    }

    return 0;
}

static int raster_sphere(const MJNode* obj, ShapeGrid* g, char* err) {
    float O[3];

    if (get_triplet(mj_get(obj, "O"), O)) {
        snprintf(err, 256, "A Sphere command misses O field");
        return 1;
    }

    const MJNode* rv = mj_get(obj, "R");

    if (!rv) {
        snprintf(err, 256, "A Sphere command misses R field");
        return 2;
    }

    float R = (float)mj_num(rv, 0.0);
    float R2 = R * R;
    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);

    for (int k = 0; k < (int)g->dim->z; k++) {
        float dz = (k + 0.5f) - O[2];

        for (int j = 0; j < (int)g->dim->y; j++) {
            float dy = (j + 0.5f) - O[1];

            for (int i = 0; i < (int)g->dim->x; i++) {
                float dx = (i + 0.5f) - O[0];

                if (dx * dx + dy * dy + dz * dz <= R2) {
                    (*g->vol)[vox_index(g, i, j, k)] = tag;
                }
            }
        }
    }

    return 0;
}

static int raster_subgrid(const MJNode* obj, ShapeGrid* g, char* err) {
    float Of[3], Sf[3];

    if (get_triplet(mj_get(obj, "O"), Of)) {
        snprintf(err, 256, "A Subgrid command misses O field");
        return 1;
    }

    if (get_triplet(mj_get(obj, "Size"), Sf)) {
        snprintf(err, 256, "A Subgrid command misses Size field");
        return 2;
    }

    int O[3], S[3];

    for (int d = 0; d < 3; d++) {
        O[d] = (int)((int)Of[d] - 1 - g->orig[d]);   /* 1-based corner */
        S[d] = (int)Sf[d];
    }

    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);

    for (int k = 0; k < (int)g->dim->z; k++) {
        if (k < O[2] || k > O[2] + S[2]) {
            continue;
        }

        for (int j = 0; j < (int)g->dim->y; j++) {
            if (j < O[1] || j > O[1] + S[1]) {
                continue;
            }

            for (int i = 0; i < (int)g->dim->x; i++) {
                if (i < O[0] || i > O[0] + S[0]) {
                    continue;
                }

                (*g->vol)[vox_index(g, i, j, k)] = tag;
            }
        }
    }

    return 0;
}

static int raster_box(const MJNode* obj, ShapeGrid* g, char* err) {
    float O[3], S[3];

    if (get_triplet(mj_get(obj, "O"), O)) {
        snprintf(err, 256, "A Box command misses O field");
        return 1;
    }

    if (get_triplet(mj_get(obj, "Size"), S)) {
        snprintf(err, 256, "A Box command misses Size field");
        return 2;
    }

    for (int d = 0; d < 3; d++) {
        O[d] -= g->orig[d];
    }

    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);

    for (int k = 0; k < (int)g->dim->z; k++) {
        float cz = k + 0.5f;

        if (cz < O[2] || cz > O[2] + S[2]) {
            continue;
        }

        for (int j = 0; j < (int)g->dim->y; j++) {
            float cy = j + 0.5f;

            if (cy < O[1] || cy > O[1] + S[1]) {
                continue;
            }

            for (int i = 0; i < (int)g->dim->x; i++) {
                float cx = i + 0.5f;

                if (cx < O[0] || cx > O[0] + S[0]) {
                    continue;
                }

                (*g->vol)[vox_index(g, i, j, k)] = tag;
            }
        }
    }

    return 0;
}

static int raster_cylinder(const MJNode* obj, ShapeGrid* g, char* err) {
    float C0[3], C1[3], v[3];

    if (get_triplet(mj_get(obj, "C0"), C0) || get_triplet(mj_get(obj, "C1"), C1)) {
        snprintf(err, 256, "A Cylinder command misses C0 or C1 field");
        return 1;
    }

    for (int d = 0; d < 3; d++) {
        C0[d] -= g->orig[d];
        C1[d] -= g->orig[d];
        v[d] = C1[d] - C0[d];
    }

    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);

    if (len == 0.f) {
        snprintf(err, 256, "Coincident end points in the definition of Cylinder command");
        return 1;
    }

    for (int d = 0; d < 3; d++) {
        v[d] /= len;
    }

    const MJNode* rv = mj_get(obj, "R");

    if (!rv) {
        snprintf(err, 256, "A Cylinder command misses R field");
        return 2;
    }

    float R = (float)mj_num(rv, 0.0);
    float R2 = R * R;
    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);

    for (int k = 0; k < (int)g->dim->z; k++) {
        float dz = (k + 0.5f) - C0[2];

        for (int j = 0; j < (int)g->dim->y; j++) {
            float dy = (j + 0.5f) - C0[1];

            for (int i = 0; i < (int)g->dim->x; i++) {
                float dx = (i + 0.5f) - C0[0];
                float along = v[0] * dx + v[1] * dy + v[2] * dz;

                if (along > len || along < 0.f) {
                    continue;
                }

                if (dx * dx + dy * dy + dz * dz - along * along <= R2) {
                    (*g->vol)[vox_index(g, i, j, k)] = tag;
                }
            }
        }
    }

    return 0;
}

static int raster_upperspace(const MJNode* obj, ShapeGrid* g, char* err) {
    const MJNode* c = mj_get(obj, "Coef");

    if (!c || c->type != MJ_ARRAY || c->count != 4) {
        snprintf(err, 256, "An UpperSpace command misses Coef field");
        return 1;
    }

    float C[4];

    for (int i = 0; i < 4; i++) {
        C[i] = (float)mj_num(mj_at(c, i), 0.0);
    }

    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);

    for (int k = 0; k < (int)g->dim->z; k++) {
        for (int j = 0; j < (int)g->dim->y; j++) {
            for (int i = 0; i < (int)g->dim->x; i++) {
                if (C[0] * (i + 0.5f) + C[1] * (j + 0.5f) + C[2] * (k + 0.5f) > C[3]) {
                    (*g->vol)[vox_index(g, i, j, k)] = tag;
                }
            }
        }
    }

    return 0;
}

static void fill_range(ShapeGrid* g, int axis, int lo, int hi, unsigned int tag) {
    unsigned int n[3] = {g->dim->x, g->dim->y, g->dim->z};

    if (lo < 0) {
        lo = 0;
    }

    if (hi > (int)n[axis]) {
        hi = (int)n[axis];
    }

    for (int k = 0; k < (int)n[2]; k++) {
        if (axis == 2 && (k < lo || k >= hi)) {
            continue;
        }

        for (int j = 0; j < (int)n[1]; j++) {
            if (axis == 1 && (j < lo || j >= hi)) {
                continue;
            }

            for (int i = 0; i < (int)n[0]; i++) {
                if (axis == 0 && (i < lo || i >= hi)) {
                    continue;
                }

                (*g->vol)[vox_index(g, i, j, k)] = tag;
            }
        }
    }
}

static int raster_layers(const MJNode* obj, ShapeGrid* g, int axis, char* err) {
    if (!obj || obj->type != MJ_ARRAY) {
        snprintf(err, 256, "A %s object must be an array", obj ? obj->key : "Layers");
        return 1;
    }

    if (obj->count == 0) {
        return 0;
    }

    const MJNode* item = obj->child;
    int single = (obj->count == 3 && item->type != MJ_ARRAY);
    int num = single ? 1 : obj->count;

    for (int i = 0; i < num; i++) {
        const MJNode* t = single ? obj : item;

        if (!t || t->count != 3) {
            snprintf(err, 256, "The %s must contain integer triplets", obj->key ? obj->key : "Layers");
            return 2;
        }

        int a = mj_int(mj_at(t, 0), 1);
        int b = mj_int(mj_at(t, 1), 1);
        unsigned int tag = (unsigned int)mj_int(mj_at(t, 2), 0);
        int lo = (a > 1 ? a : 1) - 1;
        int hi = b;

        if (hi < lo) {
            int tmp = hi;
            hi = lo;
            lo = tmp;
        }

        fill_range(g, axis, lo, hi, tag);

        if (!single) {
            item = item->next;
        }
    }

    return 0;
}

static int raster_slabs(const MJNode* obj, ShapeGrid* g, int axis, char* err) {
    const MJNode* bd = mj_get(obj, "Bound");

    if (!bd || bd->type != MJ_ARRAY) {
        snprintf(err, 256, "A %s command misses Bound field", obj->key ? obj->key : "Slabs");
        return 1;
    }

    if (bd->count == 0) {
        return 0;
    }

    unsigned int tag = (unsigned int)mj_int(mj_get(obj, "Tag"), 0);
    int single = (bd->child->type != MJ_ARRAY);
    int num = single ? 1 : bd->count;
    const MJNode* item = bd->child;
    unsigned int n[3] = {g->dim->x, g->dim->y, g->dim->z};

    for (int i = 0; i < num; i++) {
        const MJNode* pr = single ? bd : item;

        if (!pr || pr->count != 2) {
            snprintf(err, 256, "The Bound field must contain number pairs");
            return 2;
        }

        float lo = (float)mj_num(mj_at(pr, 0), 0) - 0.5f;
        float hi = (float)mj_num(mj_at(pr, 1), 0);
        lo = lo > 0.f ? lo : 0.f;
        hi = hi < (float)n[axis] ? hi : (float)n[axis];

        if (hi < lo) {
            float tmp = hi;
            hi = lo;
            lo = tmp;
        }

        fill_range(g, axis, (int)lo, (int)hi, tag);

        if (!single) {
            item = item->next;
        }
    }

    return 0;
}

static int raster_jdata(const MJNode* shapes, ShapeGrid* g, char* err) {
    /* uncompressed JData array: {"_ArraySize_":[nx,ny,nz], "_ArrayData_":[...]} */
    const MJNode* sz = mj_get(shapes, "_ArraySize_");
    const MJNode* data = mj_get(shapes, "_ArrayData_");
    float d[3];

    if (get_triplet(sz, d) || !data || data->type != MJ_ARRAY) {
        snprintf(err, 256, "Unsupported JData volume: only uncompressed _ArrayData_ is read");
        return 3;
    }

    g->dim->x = (unsigned int)d[0];
    g->dim->y = (unsigned int)d[1];
    g->dim->z = (unsigned int)d[2];
    size_t n = (size_t)g->dim->x * g->dim->y * g->dim->z;

    if ((size_t)data->count != n) {
        snprintf(err, 256, "JData volume length does not match _ArraySize_");
        return 3;
    }

    free(*g->vol);
    *g->vol = (unsigned int*)calloc(n, sizeof(unsigned int));
    const MJNode* v = data->child;

    /* JData arrays are row-major: index (i,j,k) -> ((i*ny)+j)*nz+k */
    for (size_t i = 0; i < g->dim->x; i++)
        for (size_t j = 0; j < g->dim->y; j++)
            for (size_t k = 0; k < g->dim->z; k++) {
                (*g->vol)[k * g->dim->x * g->dim->y + j * g->dim->x + i] = (unsigned int)mj_int(v, 0);
                v = v->next;
            }

    return 0;
}

int mcx_rasterize_shapes(const void* shapesnode, unsigned int** vol, MCXUInt4* dim, int issrcfrom0, int rowmajor,
                         char* err) {
    const MJNode* shapes = (const MJNode*)shapesnode;
    ShapeGrid g;
    g.vol = vol;
    g.dim = dim;
    g.rowmajor = rowmajor;

    for (int d = 0; d < 3; d++) {
        g.orig[d] = issrcfrom0 ? 0.f : 1.f;
    }

    if (shapes && shapes->type == MJ_OBJECT && mj_get(shapes, "Shapes")) {
        shapes = mj_get(shapes, "Shapes");
    }

    if (shapes && shapes->type == MJ_OBJECT && mj_get(shapes, "_ArraySize_")) {
        return raster_jdata(shapes, &g, err);
    }

    size_t n = (size_t)dim->x * dim->y * dim->z;
    free(*vol);
    *vol = NULL;

    if (n > 0) {
        *vol = (unsigned int*)malloc(n * sizeof(unsigned int));

        for (size_t i = 0; i < n; i++) {
            (*vol)[i] = 1;
        }
    }

    if (!shapes || shapes->type != MJ_ARRAY) {
        return 0;
    }

    int objcount = 1;

    for (const MJNode* el = shapes->child; el && el->child; el = el->next, objcount++) {
        const MJNode* obj = el->child;
        const char* name = obj->key ? obj->key : "";
        int status = 0;

        if (!strcmp(name, "Name")) {
            continue;
        } else if (!strcmp(name, "Origin")) {
            status = raster_origin(obj, &g, err);
        } else if (!strcmp(name, "Grid")) {
            status = raster_grid(obj, &g, err);
        } else if (!strcmp(name, "Subgrid")) {
            status = raster_subgrid(obj, &g, err);
        } else if (!strcmp(name, "Sphere")) {
            status = raster_sphere(obj, &g, err);
        } else if (!strcmp(name, "Box")) {
            status = raster_box(obj, &g, err);
        } else if (!strcmp(name, "XSlabs") || !strcmp(name, "YSlabs") || !strcmp(name, "ZSlabs")) {
            status = raster_slabs(obj, &g, name[0] - 'X', err);
        } else if (!strcmp(name, "XLayers") || !strcmp(name, "YLayers") || !strcmp(name, "ZLayers")) {
            status = raster_layers(obj, &g, name[0] - 'X', err);
        } else if (!strcmp(name, "Cylinder")) {
            status = raster_cylinder(obj, &g, err);
        } else if (!strcmp(name, "UpperSpace")) {
            status = raster_upperspace(obj, &g, err);
        } else {
            snprintf(err, 256, "The #%d element in the Shapes section has an undefined tag %s", objcount, name);
            return -(objcount + 100);
        }

        if (status) {
            return status;
        }

        if (!*vol && (dim->x * dim->y * dim->z) > 0 && strcmp(name, "Origin")) {
            snprintf(err, 256, "Shapes require a domain size (Domain.Dim or a Grid object)");
            return 4;
        }
    }

    return 0;
}
