/*
 * mcx_jdata.c - JData array codec (reference src/mcx_utils.c:4266-4478, which
 * delegates the byte codecs to the vendored zmat library).  Here zlib provides
 * both the zlib and the gzip wrapper; base64 is written from RFC 4648.  The
 * lzip/lzma/lz4 codecs of zmat are not built.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "mcx_jdata.h"

static const char* zipnames[] = {"zlib", "gzip", "base64", "lzip", "lzma", "lz4", "lz4hc"};

int mcx_zipid(const char* name) {
    for (int i = 0; i < (int)(sizeof(zipnames) / sizeof(zipnames[0])); i++) {
        if (name && !strcmp(name, zipnames[i])) {
            return i;
        }
    }

    return -1;
}

const char* mcx_zipname(int zipid) {
    return (zipid >= 0 && zipid <= mzLz4hc) ? zipnames[zipid] : "";
}

unsigned char* mcx_zip(const void* data, size_t bytes, int zipid, size_t* outbytes, const char** err) {
    *outbytes = 0;

    if (zipid == mzBase64) {
        unsigned char* z = (unsigned char*)malloc(bytes ? bytes : 1);

        if (!z) {
            *err = "out of memory";
            return NULL;
        }

        memcpy(z, data, bytes);
        *outbytes = bytes;
        return z;
    }

    if (zipid != mzZlib && zipid != mzGzip) {
        *err = "only zlib, gzip and base64 JData compression are built";
        return NULL;
    }

    z_stream zs;
    memset(&zs, 0, sizeof(zs));

    /* window bits 15: zlib wrapper; 15+16: gzip wrapper */
    if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, zipid == mzGzip ? 31 : 15, 8,
                     Z_DEFAULT_STRATEGY) != Z_OK) {
        *err = "zlib initialisation failed";
        return NULL;
    }

    size_t cap = deflateBound(&zs, (uLong)bytes) + 32;
    unsigned char* z = (unsigned char*)malloc(cap);

    if (!z) {
        deflateEnd(&zs);
        *err = "out of memory";
        return NULL;
    }

    zs.next_in = (Bytef*)data;
    zs.avail_in = (uInt)bytes;
    zs.next_out = z;
    zs.avail_out = (uInt)cap;
    int rc = deflate(&zs, Z_FINISH);
    *outbytes = cap - zs.avail_out;
    deflateEnd(&zs);

    if (rc != Z_STREAM_END) {
        free(z);
        *err = "zlib compression failed";
        return NULL;
    }

    return z;
}

unsigned char* mcx_unzip(const unsigned char* z, size_t n, int zipid, size_t expect, size_t* outbytes,
                         const char** err) {
    *outbytes = 0;

    if (zipid == mzBase64) {
        unsigned char* out = (unsigned char*)malloc(n ? n : 1);

        if (!out) {
            *err = "out of memory";
            return NULL;
        }

        memcpy(out, z, n);
        *outbytes = n;
        return out;
    }

    if (zipid != mzZlib && zipid != mzGzip) {
        *err = "only zlib, gzip and base64 JData compression are built";
        return NULL;
    }

    z_stream zs;
    memset(&zs, 0, sizeof(zs));

    /* 15+32: accept a zlib or a gzip wrapper */
    if (inflateInit2(&zs, 15 + 32) != Z_OK) {
        *err = "zlib initialisation failed";
        return NULL;
    }

    size_t cap = expect ? expect : (n * 4 + 64);
    unsigned char* out = (unsigned char*)malloc(cap ? cap : 1);

    if (!out) {
        inflateEnd(&zs);
        *err = "out of memory";
        return NULL;
    }

    zs.next_in = (Bytef*)z;
    zs.avail_in = (uInt)n;
    int rc;

    for (;;) {
        if (zs.total_out == cap) {
            size_t ncap = cap * 2;
            unsigned char* grown = (unsigned char*)realloc(out, ncap);

            if (!grown) {
                free(out);
                inflateEnd(&zs);
                *err = "out of memory";
                return NULL;
            }

            out = grown;
            cap = ncap;
        }

        zs.next_out = out + zs.total_out;
        zs.avail_out = (uInt)(cap - zs.total_out);
        rc = inflate(&zs, Z_NO_FLUSH);

        if (rc == Z_STREAM_END) {
            break;
        }

        if (rc == Z_OK || (rc == Z_BUF_ERROR && zs.avail_out == 0)) {
            continue;   /* progress made, or only the output space ran out */
        }

        break;          /* corrupted stream, or the input ended before the stream did */
    }

    *outbytes = zs.total_out;
    inflateEnd(&zs);

    if (rc != Z_STREAM_END) {
        free(out);
        *outbytes = 0;
        *err = "corrupted compressed JData stream";
        return NULL;
    }

    return out;
}

static const char b64tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

char* mcx_base64_encode(const unsigned char* z, size_t n) {
    char* out = (char*)malloc(4 * ((n + 2) / 3) + 1);

    if (!out) {
        return NULL;
    }

    size_t o = 0;

    for (size_t i = 0; i < n; i += 3) {
        size_t k = n - i;
        uint32_t v = ((uint32_t)z[i] << 16) | (k > 1 ? (uint32_t)z[i + 1] << 8 : 0) | (k > 2 ? z[i + 2] : 0);
        out[o++] = b64tab[(v >> 18) & 63];
        out[o++] = b64tab[(v >> 12) & 63];
        out[o++] = k > 1 ? b64tab[(v >> 6) & 63] : '=';
        out[o++] = k > 2 ? b64tab[v & 63] : '=';
    }

    out[o] = '\0';
    return out;
}

/* RFC 4648 decoding; characters outside the alphabet (line breaks) are skipped */
unsigned char* mcx_base64_decode(const char* text, size_t* outbytes) {
    signed char rev[256];
    memset(rev, -1, sizeof(rev));

    for (int i = 0; i < 64; i++) {
        rev[(unsigned char)b64tab[i]] = (signed char)i;
    }

    size_t n = strlen(text);
    unsigned char* out = (unsigned char*)malloc(n / 4 * 3 + 3);

    if (!out) {
        return NULL;
    }

    uint32_t acc = 0;
    int bits = 0;
    size_t o = 0;

    for (size_t i = 0; i < n && text[i] != '='; i++) {
        int v = rev[(unsigned char)text[i]];

        if (v < 0) {
            continue;
        }

        acc = (acc << 6) | (uint32_t)v;
        bits += 6;

        if (bits >= 8) {
            bits -= 8;
            out[o++] = (unsigned char)(acc >> bits);
        }
    }

    *outbytes = o;
    return out;
}

int mcx_jdata_typebytes(const char* type) {
    static const struct {
        const char* name;
        int bytes;
    } t[] = {{"int8", 1}, {"uint8", 1}, {"int16", 2}, {"uint16", 2}, {"int32", 4}, {"uint32", 4},
        {"int64", 8}, {"uint64", 8}, {"single", 4}, {"float32", 4}, {"double", 8}, {"float64", 8}
    };

    for (size_t i = 0; type && i < sizeof(t) / sizeof(t[0]); i++) {
        if (!strcmp(type, t[i].name)) {
            return t[i].bytes;
        }
    }

    return 0;
}

static void store_number(void* buf, size_t i, const char* type, double v) {
    if (!strcmp(type, "int8")) {
        ((int8_t*)buf)[i] = (int8_t)v;
    } else if (!strcmp(type, "uint8")) {
        ((uint8_t*)buf)[i] = (uint8_t)v;
    } else if (!strcmp(type, "int16")) {
        ((int16_t*)buf)[i] = (int16_t)v;
    } else if (!strcmp(type, "uint16")) {
        ((uint16_t*)buf)[i] = (uint16_t)v;
    } else if (!strcmp(type, "int32")) {
        ((int32_t*)buf)[i] = (int32_t)v;
    } else if (!strcmp(type, "uint32")) {
        ((uint32_t*)buf)[i] = (uint32_t)v;
    } else if (!strcmp(type, "int64")) {
        ((int64_t*)buf)[i] = (int64_t)v;
    } else if (!strcmp(type, "uint64")) {
        ((uint64_t*)buf)[i] = (uint64_t)v;
    } else if (!strcmp(type, "single") || !strcmp(type, "float32")) {
        ((float*)buf)[i] = (float)v;
    } else {
        ((double*)buf)[i] = v;
    }
}

/* reference mcx_jdatadecode (src/mcx_utils.c:4266-4380) */
void* mcx_jdata_decode(const MJNode* obj, unsigned int* dims, int maxdim, int* ndim, char* type, size_t typelen,
                       size_t* nelem, const char** err) {
    const MJNode* vsize = mj_get(obj, "_ArraySize_");
    const char* vtype = mj_str(mj_get(obj, "_ArrayType_"), NULL);
    const MJNode* vdata = mj_get(obj, "_ArrayData_");
    const MJNode* ztype = NULL;
    size_t n = 1;

    if (!vdata) {
        ztype = mj_get(obj, "_ArrayZipType_");
        vdata = mj_get(obj, "_ArrayZipData_");
    }

    if (!vtype || !mcx_jdata_typebytes(vtype)) {
        *err = "JData array with a missing or unsupported _ArrayType_";
        return NULL;
    }

    if (!vdata) {
        *err = "No _ArrayZipData_ or _ArrayData_ field is found";
        return NULL;
    }

    snprintf(type, typelen, "%s", vtype);
    *ndim = 0;

    if (vsize && vsize->type == MJ_ARRAY) {
        *ndim = vsize->count;

        for (const MJNode* d = vsize->child; d; d = d->next) {
            n *= (size_t)mj_num(d, 0);
        }

        for (int i = 0; i < maxdim; i++) {
            dims[i] = (i < *ndim) ? (unsigned int)mj_num(mj_at(vsize, i), 1) : 1u;
        }
    } else if (vsize && vsize->type == MJ_NUMBER) {
        *ndim = 1;
        n = (size_t)vsize->num;
        dims[0] = (unsigned int)n;

        for (int i = 1; i < maxdim; i++) {
            dims[i] = 1;
        }
    } else {
        *err = "JData array without _ArraySize_";
        return NULL;
    }

    int eb = mcx_jdata_typebytes(vtype);
    *nelem = n;

    if (ztype) {
        int zipid = mcx_zipid(mj_str(ztype, ""));
        size_t zbytes = 0, bytes = 0;

        if (zipid < 0 || vdata->type != MJ_STRING) {
            *err = "unknown _ArrayZipType_ or a non-string _ArrayZipData_";
            return NULL;
        }

        unsigned char* z = mcx_base64_decode(vdata->str, &zbytes);

        if (!z) {
            *err = "out of memory";
            return NULL;
        }

        unsigned char* raw = mcx_unzip(z, zbytes, zipid, n * eb, &bytes, err);
        free(z);

        if (raw && bytes != n * eb) {
            free(raw);
            *err = "decoded JData size does not match _ArraySize_";
            return NULL;
        }

        return raw;
    }

    if (vdata->type != MJ_ARRAY) {
        *err = "JData array constructs are not supported";
        return NULL;
    }

    /* _ArrayData_: flat numbers, or nested rows flattened in order */
    void* buf = calloc(n ? n : 1, eb);
    size_t i = 0;
    const MJNode* stack[16];
    int depth = 0;
    const MJNode* cur = vdata->child;

    if (!buf) {
        *err = "out of memory";
        return NULL;
    }

    while (cur || depth > 0) {
        if (!cur) {
            cur = stack[--depth];
            continue;
        }

        if (cur->type == MJ_ARRAY) {
            if (depth == 16) {
                free(buf);
                *err = "JData array nested too deeply";
                return NULL;
            }

            stack[depth++] = cur->next;
            cur = cur->child;
            continue;
        }

        if (cur->type != MJ_NUMBER || i >= n) {
            free(buf);
            *err = "JData _ArrayData_ does not match _ArraySize_";
            return NULL;
        }

        store_number(buf, i++, vtype, cur->num);
        cur = cur->next;
    }

    if (i != n) {
        free(buf);
        *err = "JData _ArrayData_ does not match _ArraySize_";
        return NULL;
    }

    return buf;
}
