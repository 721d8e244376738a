/*
 * mcx_output.c - volumetric and detected-photon output files.
 *
 * Formats and entry points follow the reference's writers (src/mcx_utils.c):
 *   mcx_savedata       dispatch on cfg->outputformat                      :886-950
 *     .mc2 / .ubj      raw float32, x fastest
 *     .tx3             GL_RGBA32F tag, 3 dims, float32 data
 *   mcx_savenii        .nii (348-byte NIfTI-1 header, 4-byte extender, data)
 *                      or .hdr + .img                                     :487-584
 *   mcx_savejnii       .jnii: JSON JNIfTI, data compressed + base64       :761-880
 *   mcx_savebnii       .bnii: the same document in BJData (binary JSON)   :598-747
 *   mcx_savedetphoton  .mch: History header + records (+ seeds)           :963-998
 *   mcx_savejdet       <session>_detp.jdat: JSON MCXData                   :1010-1182
 * The NIfTI-1 header is written field by field from the published NIfTI-1
 * layout; BJData follows the NeuroJSON BJData Draft 2 specification
 * (little-endian, typed and counted containers).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mcx_jdata.h"
#include "mcx_utils.h"

#define NII_HDR_BYTES 348
#define NII_UNITS_MM 2
#define NII_UNITS_USEC 24
#define NII_TYPE_UINT16 512
#define NII_TYPE_FLOAT32 16
#define GL_RGBA32F_TAG 0x8814u

static const char* format_ext[] = {"mc2", "nii", "hdr", "ubj", "tx3", "jnii", "bnii"};

static const char* output_desc[] = {"Fluence rate (W/mm^2)", "Fluence (J/mm^2)", "Voxel-wise energy deposit (J)",
                                    "Jacobian for mua (J/mm)", "Scattering count", "Partial momentum transfer"
                                   };

/* <rootpath>/<session>, or <session> without a root path */
static void output_name(const Config* cfg, char* name, size_t len) {
    if (cfg->rootpath[0]) {
        snprintf(name, len, "%s/%s", cfg->rootpath, cfg->session);
    } else {
        snprintf(name, len, "%s", cfg->session);
    }
}

static FILE* open_file(const char* name, const char* ext, const char* mode) {
    char fname[MAX_FULL_PATH + 32];
    snprintf(fname, sizeof(fname), "%s%s", name, ext);
    FILE* fp = fopen(fname, mode);

    if (!fp) {
        MCX_ERROR(-2, "can not save data to disk");
    }

    return fp;
}

static void write_all(FILE* fp, const void* p, size_t size, size_t n) {
    if (n && fwrite(p, size, n, fp) != n) {
        fclose(fp);
        MCX_ERROR(-2, "incomplete write of an output file");
    }
}

static void check_zipid(const Config* cfg) {
    if (cfg->zipid != mzZlib && cfg->zipid != mzGzip && cfg->zipid != mzBase64) {
        MCX_ERROR(-2, "only zlib, gzip and base64 JData compression are built");
    }
}

static const char* describe(const Config* cfg) {
    return (cfg->outputtype >= 0 && cfg->outputtype < 6) ? output_desc[(int)cfg->outputtype] : NULL;
}

/* compress by cfg->zipid, reporting the ratios as the reference does
   (src/mcx_utils.c:4410-4440); the result is malloc'ed */
static unsigned char* compress_payload(const Config* cfg, const void* data, size_t bytes, size_t* zbytes) {
    FILE* flog = cfg->flog ? cfg->flog : stdout;
    const char* err = NULL;

    if (!cfg->isdumpjson) {
        fprintf(flog, "compressing data [%s] ...", mcx_zipname(cfg->zipid));
    }

    unsigned char* z = mcx_zip(data, bytes, cfg->zipid, zbytes, &err);

    if (!z) {
        MCX_ERROR(-2, err);
    }

    if (!cfg->isdumpjson) {
        fprintf(flog, "compression ratio: %.1f%%\t", bytes ? *zbytes * 100.f / bytes : 0.f);
    }

    return z;
}

/* ---------------- NIfTI-1 / Analyze (reference mcx_savenii) ---------------- */
static void put(unsigned char* h, size_t off, const void* v, size_t n) {
    memcpy(h + off, v, n);
}

static void put_i16(unsigned char* h, size_t off, int v) {
    short s = (short)v;
    put(h, off, &s, 2);
}

static void put_f32(unsigned char* h, size_t off, float v) {
    put(h, off, &v, 4);
}

void mcx_savenii(float* dat, size_t len, char* name, int type32bit, int outputformatid, Config* cfg) {
    unsigned char h[NII_HDR_BYTES];
    size_t nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    const void* payload = dat;
    unsigned short* mask = NULL;
    size_t nelem = len, elem = 4;
    int32_t sizeof_hdr = NII_HDR_BYTES;

    if (outputformatid != ofNifti && outputformatid != ofAnalyze) {
        MCX_ERROR(-9, "Output format is not supported");
    }

    /* field offsets of the published nifti1.h header */
    memset(h, 0, sizeof(h));
    put(h, 0, &sizeof_hdr, 4);
    put_i16(h, 40, 4);                                   /* dim[0] */
    put_i16(h, 42, (int)cfg->dim.x);
    put_i16(h, 44, (int)cfg->dim.y);
    put_i16(h, 46, (int)cfg->dim.z);
    put_i16(h, 48, (int)(nvox ? len / nvox : 0));
    put_i16(h, 68, 0);                                   /* intent_code: none */
    put_i16(h, 70, type32bit);                           /* datatype */
    put_i16(h, 72, 32);                                  /* bitpix */
    put_f32(h, 80, cfg->unitinmm);                       /* pixdim[1..3] */
    put_f32(h, 84, cfg->unitinmm);
    put_f32(h, 88, cfg->unitinmm);

    if (type32bit == NII_TYPE_FLOAT32) {
        put_f32(h, 92, cfg->tstep * 1e6f);               /* pixdim[4]: microseconds */
    } else {
        /* a uint32 label volume becomes {label, detector bit} uint16 pairs */
        mask = (unsigned short*)malloc((len ? len : 1) * 2 * sizeof(unsigned short));

        if (!mask) {
            MCX_ERROR(-9, "out of memory");
        }

        for (size_t i = 0; i < len; i++) {
            unsigned int v = ((const unsigned int*)dat)[i];
            mask[i * 2] = (unsigned short)(v & MED_MASK);
            mask[i * 2 + 1] = (unsigned short)((v & DET_MASK) >> 31);
        }

        put_i16(h, 70, NII_TYPE_UINT16);
        put_i16(h, 72, 16);
        put_i16(h, 42, 2);
        put_i16(h, 44, (int)cfg->dim.x);
        put_i16(h, 46, (int)cfg->dim.y);
        put_i16(h, 48, (int)cfg->dim.z);
        put_f32(h, 80, 1.f);
        put_f32(h, 92, cfg->unitinmm);
        payload = mask;
        nelem = len * 2;
        elem = 2;
    }

    int single = (outputformatid == ofNifti);
    put_f32(h, 108, single ? (float)(NII_HDR_BYTES + 4) : 0.f);   /* vox_offset */
    put_f32(h, 112, 0.f);                                          /* scl_slope */
    h[123] = NII_UNITS_MM | NII_UNITS_USEC;                        /* xyzt_units */
    put(h, 344, single ? "n+1" : "ni1", 4);                        /* magic, NUL included */

    FILE* fp = open_file(name, single ? ".nii" : ".hdr", "wb");
    write_all(fp, h, 1, NII_HDR_BYTES);

    if (single) {
        const unsigned char extender[4] = {0, 0, 0, 0};
        write_all(fp, extender, 1, 4);
    } else {
        fclose(fp);
        fp = open_file(name, ".img", "wb");
    }

    /* the reference writes an unfilled buffer in place of float data
       (src/mcx_utils.c:505-506, :550-553); the data itself is written here */
    write_all(fp, payload, elem, nelem);
    fclose(fp);
    free(mask);
}

/* ---------------- JNIfTI, text form (reference mcx_savejnii) ---------------- */
static void json_string(FILE* fp, const char* s) {
    fputc('"', fp);

    for (; *s; s++) {
        unsigned char c = (unsigned char)*s;

        if (c == '"' || c == '\\') {
            fprintf(fp, "\\%c", c);
        } else if (c < 0x20) {
            fprintf(fp, "\\u%04x", c);
        } else {
            fputc(c, fp);
        }
    }

    fputc('"', fp);
}

/* reference mcx_jdataencode (src/mcx_utils.c:4396-4478), text form */
static void json_jdata(FILE* fp, const Config* cfg, const char* type, const unsigned int* dims, int ndim,
                       int iscol, const void* data, size_t bytes, const char* indent) {
    size_t n = 1, zbytes = 0;

    for (int i = 0; i < ndim; i++) {
        n *= dims[i];
    }

    unsigned char* z = compress_payload(cfg, data, bytes, &zbytes);
    char* b64 = mcx_base64_encode(z, zbytes);
    free(z);

    if (!b64) {
        fclose(fp);
        MCX_ERROR(-2, "out of memory");
    }

    if (!cfg->isdumpjson) {
        fprintf(cfg->flog ? cfg->flog : stdout, "after encoding: %.1f%%\n", bytes ? strlen(b64) * 100.f / bytes : 0.f);
    }

    fprintf(fp, "{\n%s\t\"_ArrayType_\":\t\"%s\",\n%s\t\"_ArraySize_\":\t[", indent, type, indent);

    for (int i = 0; i < ndim; i++) {
        fprintf(fp, "%s%u", i ? ", " : "", dims[i]);
    }

    fprintf(fp, "],\n");

    if (iscol) {
        fprintf(fp, "%s\t\"_ArrayOrder_\":\t\"c\",\n", indent);
    }

    fprintf(fp, "%s\t\"_ArrayZipType_\":\t\"%s\",\n%s\t\"_ArrayZipSize_\":\t%zu,\n%s\t\"_ArrayZipData_\":\t\"",
            indent, mcx_zipname(cfg->zipid), indent, n, indent);
    fputs(b64, fp);
    fprintf(fp, "\"\n%s}", indent);
    free(b64);
}

void mcx_savejnii(float* vol, int ndim, unsigned int* dims, float* voxelsize, char* name, int isfloat, int iscol,
                  Config* cfg) {
    size_t n = 1;
    const char* desc = describe(cfg);
    check_zipid(cfg);

    for (int i = 0; i < ndim; i++) {
        n *= dims[i];
    }

    FILE* fp = open_file(name, ".jnii", "wt");
    fprintf(fp, "{\n\t\"_DataInfo_\":\t{\n\t\t\"JNIFTIVersion\":\t\"0.5\",\n\t\t\"Comment\":\t\"Created by mcxcl_amd\",\n"
            "\t\t\"AnnotationFormat\":\t\"https://neurojson.org/jnifti/draft1\",\n"
            "\t\t\"SerialFormat\":\t\"https://json.org\"\n\t},\n");
    fprintf(fp, "\t\"NIFTIHeader\":\t{\n\t\t\"NIIHeaderSize\":\t348,\n\t\t\"Dim\":\t[");

    for (int i = 0; i < ndim; i++) {
        fprintf(fp, "%s%u", i ? ", " : "", dims[i]);
    }

    fprintf(fp, "],\n\t\t\"Param1\":\t0,\n\t\t\"Param2\":\t0,\n\t\t\"Param3\":\t0,\n\t\t\"Intent\":\t0,\n"
            "\t\t\"DataType\":\t\"%s\",\n\t\t\"BitDepth\":\t32,\n\t\t\"FirstSliceID\":\t0,\n\t\t\"VoxelSize\":\t[",
            isfloat ? "single" : "uint32");

    for (int i = 0; i < ndim; i++) {
        fprintf(fp, "%s%.9g", i ? ", " : "", voxelsize[i]);
    }

    fprintf(fp, "],\n\t\t\"Orientation\":\t{\n\t\t\t\"x\":\t\"r\",\n\t\t\t\"y\":\t\"a\",\n\t\t\t\"z\":\t\"s\"\n\t\t},\n"
            "\t\t\"ScaleSlope\":\t1,\n\t\t\"ScaleOffset\":\t0,\n\t\t\"LastSliceID\":\t%u,\n\t\t\"SliceType\":\t1,\n"
            "\t\t\"Unit\":\t{\n\t\t\t\"L\":\t\"mm\",\n\t\t\t\"T\":\t\"s\"\n\t\t},\n\t\t\"MaxIntensity\":\t1,\n"
            "\t\t\"MinIntensity\":\t0,\n\t\t\"SliceTime\":\t0,\n\t\t\"TimeOffset\":\t0,\n"
            "\t\t\"Description\":\t\"MCX volumetric output%s%s\",\n\t\t\"AuxFile\":\t\"\",\n\t\t\"QForm\":\t0,\n"
            "\t\t\"SForm\":\t1,\n\t\t\"Quatern\":\t{\n\t\t\t\"b\":\t0,\n\t\t\t\"c\":\t0,\n\t\t\t\"d\":\t0\n\t\t},\n"
            "\t\t\"QuaternOffset\":\t{\n\t\t\t\"x\":\t0,\n\t\t\t\"y\":\t0,\n\t\t\t\"z\":\t0\n\t\t},\n"
            "\t\t\"Affine\":\t[[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]],\n\t\t\"Name\":\t",
            cfg->maxgate, desc ? ": " : "", desc ? desc : "");
    json_string(fp, cfg->session);
    fprintf(fp, ",\n\t\t\"NIIFormat\":\t\"jnifti\"\n\t},\n\t\"NIFTIData\":\t");
    json_jdata(fp, cfg, isfloat ? "single" : "uint32", dims, ndim, iscol, vol, n * 4, "\t");
    fputs("\n}\n", fp);
    fclose(fp);
}

/* ---------------- JNIfTI, BJData form (reference mcx_savebnii) ---------------- */
typedef struct {
    unsigned char* p;
    size_t n, cap;
} BJBuf;

static void bj_raw(BJBuf* b, const void* v, size_t n) {
    if (b->n + n > b->cap) {
        size_t cap = (b->cap ? b->cap * 2 : 4096);

        while (cap < b->n + n) {
            cap *= 2;
        }

        unsigned char* grown = (unsigned char*)realloc(b->p, cap);

        if (!grown) {
            free(b->p);
            MCX_ERROR(-2, "out of memory");
        }

        b->p = grown;
        b->cap = cap;
    }

    memcpy(b->p + b->n, v, n);
    b->n += n;
}

static void bj_mark(BJBuf* b, char m) {
    bj_raw(b, &m, 1);
}

/* an integer in the smallest BJData type that holds it (U u m M, little-endian) */
static void bj_uint(BJBuf* b, uint64_t v) {
    if (v <= 0xFFu) {
        uint8_t x = (uint8_t)v;
        bj_mark(b, 'U');
        bj_raw(b, &x, 1);
    } else if (v <= 0xFFFFu) {
        uint16_t x = (uint16_t)v;
        bj_mark(b, 'u');
        bj_raw(b, &x, 2);
    } else if (v <= 0xFFFFFFFFu) {
        uint32_t x = (uint32_t)v;
        bj_mark(b, 'm');
        bj_raw(b, &x, 4);
    } else {
        bj_mark(b, 'M');
        bj_raw(b, &v, 8);
    }
}

/* object keys and string payloads: length, then bytes (no 'S' marker) */
static void bj_name(BJBuf* b, const char* s) {
    size_t n = strlen(s);
    bj_uint(b, n);
    bj_raw(b, s, n);
}

static void bj_key_str(BJBuf* b, const char* key, const char* s) {
    bj_name(b, key);
    bj_mark(b, 'S');
    bj_name(b, s);
}

static void bj_key_uint(BJBuf* b, const char* key, uint64_t v) {
    bj_name(b, key);
    bj_uint(b, v);
}

static void bj_key_char(BJBuf* b, const char* key, char c) {
    bj_name(b, key);
    bj_mark(b, 'C');
    bj_mark(b, c);
}

/* strongly typed, counted array: [ $ type # count payload */
static void bj_typed(BJBuf* b, char type, const void* data, size_t elem, size_t count) {
    bj_mark(b, '[');
    bj_mark(b, '$');
    bj_mark(b, type);
    bj_mark(b, '#');
    bj_uint(b, count);
    bj_raw(b, data, elem * count);
}

void mcx_savebnii(float* vol, int ndim, unsigned int* dims, float* voxelsize, char* name, int isfloat, int iscol,
                  Config* cfg) {
    BJBuf b = {NULL, 0, 0};
    size_t n = 1, zbytes = 0;
    const int32_t affine[6] = {0, 0, 1, 0, 0, 0};
    const char* desc = describe(cfg);
    char text[256];
    check_zipid(cfg);

    for (int i = 0; i < ndim; i++) {
        n *= dims[i];
    }

    bj_mark(&b, '{');
    bj_name(&b, "_DataInfo_");
    bj_mark(&b, '{');
    bj_key_str(&b, "JNIFTIVersion", "0.5");
    bj_key_str(&b, "Comment", "Created by mcxcl_amd");
    bj_key_str(&b, "AnnotationFormat", "https://neurojson.org/jnifti/draft1");
    bj_key_str(&b, "SerialFormat", "https://neurojson.org/bjdata/draft2");
    bj_mark(&b, '}');

    bj_name(&b, "NIFTIHeader");
    bj_mark(&b, '{');
    bj_key_uint(&b, "NIIHeaderSize", 348);
    bj_name(&b, "Dim");
    bj_typed(&b, 'm', dims, 4, (size_t)ndim);
    bj_key_uint(&b, "Param1", 0);
    bj_key_uint(&b, "Param2", 0);
    bj_key_uint(&b, "Param3", 0);
    bj_key_uint(&b, "Intent", 0);
    bj_key_str(&b, "DataType", isfloat ? "single" : "uint32");
    bj_key_uint(&b, "BitDepth", 32);
    bj_key_uint(&b, "FirstSliceID", 0);
    bj_name(&b, "VoxelSize");
    bj_typed(&b, 'd', voxelsize, 4, (size_t)ndim);
    bj_name(&b, "Orientation");
    bj_mark(&b, '{');
    bj_key_char(&b, "x", 'r');
    bj_key_char(&b, "y", 'a');
    bj_key_char(&b, "z", 's');
    bj_mark(&b, '}');
    bj_key_uint(&b, "ScaleSlope", 1);
    bj_key_uint(&b, "ScaleOffset", 0);
    bj_key_uint(&b, "LastSliceID", cfg->maxgate);
    bj_key_uint(&b, "SliceType", 1);
    bj_name(&b, "Unit");
    bj_mark(&b, '{');
    bj_key_str(&b, "L", "mm");
    bj_key_str(&b, "T", "s");
    bj_mark(&b, '}');
    bj_key_uint(&b, "MaxIntensity", 1);
    bj_key_uint(&b, "MinIntensity", 0);
    bj_key_uint(&b, "SliceTime", 0);
    bj_key_uint(&b, "TimeOffset", 0);
    snprintf(text, sizeof(text), "MCX volumetric output%s%s", desc ? ": " : "", desc ? desc : "");
    bj_key_str(&b, "Description", text);
    bj_key_str(&b, "AuxFile", "");
    bj_key_uint(&b, "QForm", 0);
    bj_key_uint(&b, "SForm", 1);
    bj_name(&b, "Quatern");
    bj_mark(&b, '{');
    bj_key_uint(&b, "b", 0);
    bj_key_uint(&b, "c", 0);
    bj_key_uint(&b, "d", 0);
    bj_mark(&b, '}');
    bj_name(&b, "QuaternOffset");
    bj_mark(&b, '{');
    bj_key_uint(&b, "x", 0);
    bj_key_uint(&b, "y", 0);
    bj_key_uint(&b, "z", 0);
    bj_mark(&b, '}');
    bj_name(&b, "Affine");
    bj_mark(&b, '[');
    bj_typed(&b, 'l', affine + 2, 4, 4);
    bj_typed(&b, 'l', affine + 1, 4, 4);
    bj_typed(&b, 'l', affine, 4, 4);
    bj_mark(&b, ']');
    bj_key_str(&b, "Name", cfg->session);
    bj_key_str(&b, "NIIFormat", "jnifti");
    bj_mark(&b, '}');

    /* NIFTIData: the annotated array with its compressed bytes as a uint8 vector */
    bj_name(&b, "NIFTIData");
    bj_mark(&b, '{');
    bj_key_str(&b, "_ArrayType_", isfloat ? "single" : "uint32");
    bj_name(&b, "_ArraySize_");
    bj_typed(&b, 'm', dims, 4, (size_t)ndim);

    if (iscol) {
        bj_key_str(&b, "_ArrayOrder_", "c");
    }

    unsigned char* z = compress_payload(cfg, vol, n * 4, &zbytes);

    if (!cfg->isdumpjson) {
        fprintf(cfg->flog ? cfg->flog : stdout, "\n");
    }

    bj_key_str(&b, "_ArrayZipType_", mcx_zipname(cfg->zipid));
    bj_key_uint(&b, "_ArrayZipSize_", n);
    bj_name(&b, "_ArrayZipData_");
    bj_typed(&b, 'U', z, 1, zbytes);
    free(z);
    bj_mark(&b, '}');
    bj_mark(&b, '}');

    FILE* fp = open_file(name, ".bnii", "wb");
    write_all(fp, b.p, 1, b.n);
    fclose(fp);
    free(b.p);
}

/* ---------------- volumetric dispatch (reference mcx_savedata) ---------------- */
void mcx_savedata(float* dat, size_t len, Config* cfg) {
    char name[MAX_FULL_PATH];
    output_name(cfg, name, sizeof(name));

    if (cfg->outputformat == ofNifti || cfg->outputformat == ofAnalyze) {
        mcx_savenii(dat, len, name, NII_TYPE_FLOAT32, cfg->outputformat, cfg);
        return;
    }

    if (cfg->outputformat == ofJNifti || cfg->outputformat == ofBJNifti) {
        /* [x, y, z, gates, sources(, stacked sources or replay detectors)]
           (src/mcx_utils.c:901-926) */
        unsigned int dims[6] = {cfg->dim.x, cfg->dim.y, cfg->dim.z, cfg->maxgate, cfg->srcnum, 1};
        float voxelsize[6] = {cfg->steps.x, cfg->steps.y, cfg->steps.z, cfg->tstep, 1.f, 1.f};

        if (cfg->extrasrclen && cfg->srcid < 0) {
            dims[5] = cfg->extrasrclen + 1;
        }

        if (cfg->seed == SEED_FROM_FILE && cfg->replaydet == -1 && cfg->detnum > 1) {
            dims[5] *= cfg->detnum;
        }

        if (cfg->outputformat == ofJNifti) {
            mcx_savejnii(dat, 5 + (dims[5] > 1), dims, voxelsize, name, 1, 1, cfg);
        } else {
            mcx_savebnii(dat, 5 + (dims[5] > 1), dims, voxelsize, name, 1, 1, cfg);
        }

        return;
    }

    if (cfg->outputformat < 0 || cfg->outputformat > ofBJNifti) {
        MCX_ERROR(-2, "unknown output format");
    }

    char ext[8];
    snprintf(ext, sizeof(ext), ".%s", format_ext[(int)cfg->outputformat]);
    FILE* fp = open_file(name, ext, "wb");

    if (cfg->outputformat == ofTX3) {
        unsigned int tag = GL_RGBA32F_TAG;
        write_all(fp, &tag, sizeof(tag), 1);
        write_all(fp, &cfg->dim.x, sizeof(int), 3);
    }

    write_all(fp, dat, sizeof(float), len);
    fclose(fp);
}

/* ---------------- detected photons ---------------- */
static void json_info_uint(FILE* fp, const char* key, unsigned int v) {
    fprintf(fp, "\t\t\t\"%s\":\t%u,\n", key, v);
}

/* reference mcx_savejdet (src/mcx_utils.c:1010-1182): the History fields as
   MCXData.Info and one JData array per savedetflag column group.  Two points
   differ on purpose: nscat is stored in the records as raw uint32 bits
   (src/mcx_core.cl:2511), which are written as they are rather than converted
   by value, and with several sources the source index that the kernel packs
   into the high 16 bits of detid (src/mcx_core.cl:887-888) is split from detid
   rather than from nscat.  Trajectory files (detected == 0) are not produced
   by this build. */
void mcx_savejdet(float* ppath, void* seeds, unsigned int count, int doappend, Config* cfg) {
    const History* his = &cfg->his;
    char name[MAX_FULL_PATH];
    check_zipid(cfg);
    output_name(cfg, name, sizeof(name));
    FILE* fp = open_file(name, "_detp.jdat", doappend ? "at" : "wt");
    fprintf(fp, "{\n\t\"MCXData\":\t{\n\t\t\"Info\":\t{\n");
    json_info_uint(fp, "Version", his->version);
    json_info_uint(fp, "MediaNum", his->maxmedia);
    json_info_uint(fp, "DetNum", his->detnum);
    json_info_uint(fp, "ColumnNum", his->colcount);
    json_info_uint(fp, "TotalPhoton", his->totalphoton);
    json_info_uint(fp, "DetectedPhoton", count);
    json_info_uint(fp, "SavedPhoton", his->savedphoton);
    fprintf(fp, "\t\t\t\"LengthUnit\":\t%.9g,\n", his->unitinmm);
    json_info_uint(fp, "SeedByte", his->seedbyte);
    fprintf(fp, "\t\t\t\"Normalizer\":\t%.9g,\n\t\t\t\"Repeat\":\t%d,\n", his->normalizer, his->respin);
    json_info_uint(fp, "SrcNum", his->srcnum);
    json_info_uint(fp, "SaveDetFlag", his->savedetflag);
    json_info_uint(fp, "TotalSource", his->totalsource);
    fprintf(fp, "\t\t\t\"Media\":\t[");

    for (unsigned int i = 0; i < cfg->medianum; i++) {
        fprintf(fp, "%s{\n\t\t\t\t\t\"mua\":\t%.9g,\n\t\t\t\t\t\"mus\":\t%.9g,\n\t\t\t\t\t\"g\":\t%.9g,\n"
                "\t\t\t\t\t\"n\":\t%.9g\n\t\t\t\t}", i ? ", " : "", cfg->prop[i].mua / cfg->unitinmm,
                cfg->prop[i].mus / cfg->unitinmm, cfg->prop[i].g, cfg->prop[i].n);
    }

    fprintf(fp, "]\n\t\t},\n\t\t\"PhotonData\":\t{\n");
    /* column groups in savedetflag bit order (src/mcx_utils.c:1072-1074) */
    const unsigned int ncol[7] = {1, his->maxmedia, his->maxmedia, his->maxmedia, 3, 3, 1};
    static const char* type[7] = {"uint32", "uint32", "single", "single", "single", "single", "single"};
    static const char* colname[7] = {"detid", "nscat", "ppath", "mom", "p", "v", "w0"};
    size_t rows = count;
    unsigned int col = 0;
    int first = 1;

    for (int id = 0; id < 7; id++) {
        if (!((cfg->savedetflag >> id) & 1)) {
            continue;
        }

        unsigned int dims[2] = {(unsigned int)rows, ncol[id]};
        size_t n = rows * ncol[id];
        uint32_t* buf = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));

        if (!buf) {
            fclose(fp);
            MCX_ERROR(-2, "out of memory");
        }

        for (size_t i = 0; i < rows; i++) {
            for (unsigned int j = 0; j < ncol[id]; j++) {
                float v = ppath[i * his->colcount + col + j];

                if (id == 0) {
                    buf[i * ncol[id] + j] = (uint32_t)v;
                } else {
                    memcpy(buf + i * ncol[id] + j, &v, 4);
                }
            }
        }

        if (id == 0 && his->totalsource > 1) {
            uint32_t* src = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));

            if (!src) {
                free(buf);
                fclose(fp);
                MCX_ERROR(-2, "out of memory");
            }

            for (size_t i = 0; i < n; i++) {
                src[i] = buf[i] >> 16;
                buf[i] &= 0xFFFFu;
            }

            fprintf(fp, "%s\t\t\t\"srcid\":\t", first ? "" : ",\n");
            json_jdata(fp, cfg, "uint32", dims, 2, 0, src, n * 4, "\t\t\t");
            free(src);
            first = 0;
        }

        fprintf(fp, "%s\t\t\t\"%s\":\t", first ? "" : ",\n", colname[id]);
        json_jdata(fp, cfg, type[id], dims, 2, 0, buf, n * 4, "\t\t\t");
        free(buf);
        col += ncol[id];
        first = 0;
    }

    if (cfg->issaveseed && seeds && his->seedbyte) {
        unsigned int dims[2] = {(unsigned int)rows, his->seedbyte};
        fprintf(fp, "%s\t\t\t\"seed\":\t", first ? "" : ",\n");
        json_jdata(fp, cfg, "uint8", dims, 2, 0, seeds, rows * his->seedbyte, "\t\t\t");
    }

    fputs("\n\t\t}\n\t}\n}\n", fp);
    fclose(fp);
}

/* reference mcx_savedetphoton (src/mcx_utils.c:963-998): the History header as
   the caller filled it, the records and, with issaveseed, the seeds; the
   fourth argument appends to an existing file (the reference header names it
   seedbyte, src/mcx_utils.h:312, its definition uses it as doappend) */
void mcx_savedetphoton(float* ppath, void* seeds, int count, int doappend, Config* cfg) {
    if (cfg->outputformat == ofJNifti || cfg->outputformat == ofBJNifti) {
        mcx_savejdet(ppath, seeds, (unsigned int)(count > 0 ? count : 0), doappend, cfg);
        return;
    }

    char name[MAX_FULL_PATH];
    output_name(cfg, name, sizeof(name));
    FILE* fp = open_file(name, (cfg->his.detected == 0 && cfg->his.savedphoton) ? ".mct" : ".mch",
                         doappend ? "ab" : "wb");
    size_t rows = count > 0 ? (size_t)count : 0;
    write_all(fp, &cfg->his, sizeof(History), 1);
    write_all(fp, ppath, sizeof(float), rows * cfg->his.colcount);

    if (cfg->issaveseed && seeds) {
        write_all(fp, seeds, cfg->his.seedbyte, rows);
    }

    fclose(fp);
}
