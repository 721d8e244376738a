/*
 * mcx_json.h - minimal JSON reader for simulation inputs.
 *
 * The reference vendors cJSON (src/cjson) for its JSON inputs; this is a small
 * independent recursive-descent parser that keeps the document as a tree of
 * nodes with ordered children, which is what the Session/Domain/Optode/
 * Forward/Shapes readers need.
 */
#ifndef MCX_JSON_H
#define MCX_JSON_H

#ifdef __cplusplus
extern "C" {
#endif

enum MJType {MJ_NULL, MJ_FALSE, MJ_TRUE, MJ_NUMBER, MJ_STRING, MJ_ARRAY, MJ_OBJECT};

typedef struct MJNode {
    int type;
    char* key;              /* member name when the parent is an object */
    double num;             /* MJ_NUMBER value; 0/1 for booleans */
    char* str;              /* MJ_STRING value */
    struct MJNode* child;   /* first element or member */
    struct MJNode* next;    /* next sibling */
    int count;              /* number of children */
} MJNode;

MJNode* mj_parse(const char* text, char* err, int errlen);
void mj_free(MJNode* node);
MJNode* mj_get(const MJNode* obj, const char* key);
MJNode* mj_at(const MJNode* arr, int idx);
double mj_num(const MJNode* node, double fallback);
int mj_int(const MJNode* node, int fallback);
const char* mj_str(const MJNode* node, const char* fallback);
char* mj_print(const MJNode* node);   /* compact serialisation, malloc'ed */

#ifdef __cplusplus
}
#endif

#endif
