/*
 * mcx_jdata.h - JData array codec: base64, zlib/gzip and the annotated-array
 * objects ({_ArrayType_, _ArraySize_, _ArrayZipType_, _ArrayZipSize_,
 * _ArrayZipData_} or {_ArrayType_, _ArraySize_, _ArrayData_}) that the
 * reference writes and reads for .jnii volumes, .jdat detected photons and
 * JSON-embedded shapes (reference src/mcx_utils.c:4266-4478).
 */
#ifndef MCX_JDATA_H
#define MCX_JDATA_H

#include <stddef.h>

#include "mcx_json.h"

#ifdef __cplusplus
extern "C" {
#endif

enum MCXZip {mzZlib, mzGzip, mzBase64, mzLzip, mzLzma, mzLz4, mzLz4hc};

/* index of a codec name ("zlib", "gzip", "base64", ...), -1 if unknown */
int mcx_zipid(const char* name);
const char* mcx_zipname(int zipid);

/* compress (zlib or gzip wrapper) or copy (base64); malloc'ed result.
   Returns NULL with *err set for codecs that are not built. */
unsigned char* mcx_zip(const void* data, size_t bytes, int zipid, size_t* outbytes, const char** err);
/* inverse of mcx_zip; expect is the decoded byte count when known (0 if not) */
unsigned char* mcx_unzip(const unsigned char* z, size_t n, int zipid, size_t expect, size_t* outbytes,
                         const char** err);

char* mcx_base64_encode(const unsigned char* data, size_t n);
unsigned char* mcx_base64_decode(const char* text, size_t* outbytes);

/* bytes per element of a JData type name ("uint8", "single", ...), 0 if unknown */
int mcx_jdata_typebytes(const char* type);

/* decode an annotated array into a malloc'ed buffer of its declared type;
   dims receives up to maxdim sizes (row-major as the file gives them).
   Returns NULL and sets *err on malformed input. */
void* mcx_jdata_decode(const MJNode* obj, unsigned int* dims, int maxdim, int* ndim, char* type, size_t typelen,
                       size_t* nelem, const char** err);

#ifdef __cplusplus
}
#endif

#endif
