/*
 * mcx_json.c - minimal recursive-descent JSON reader (RFC 8259 subset:
 * objects, arrays, strings with escapes, numbers, true/false/null).
 */
#include "mcx_json.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char* p;
    const char* start;
    char* err;
    int errlen;
    int failed;
} MJParser;

static void mj_fail(MJParser* ps, const char* msg) {
    if (!ps->failed && ps->err && ps->errlen > 0) {
        snprintf(ps->err, ps->errlen, "JSON error at offset %ld: %s", (long)(ps->p - ps->start), msg);
    }

    ps->failed = 1;
}

static void mj_skipws(MJParser* ps) {
    while (*ps->p && isspace((unsigned char)*ps->p)) {
        ps->p++;
    }
}

static MJNode* mj_new(int type) {
    MJNode* n = (MJNode*)calloc(1, sizeof(MJNode));
    n->type = type;
    return n;
}

static MJNode* mj_value(MJParser* ps);

static void mj_put_utf8(char** out, unsigned cp) {
    char* o = *out;

    if (cp < 0x80) {
        *o++ = (char)cp;
    } else if (cp < 0x800) {
        *o++ = (char)(0xC0 | (cp >> 6));
        *o++ = (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
        *o++ = (char)(0xE0 | (cp >> 12));
        *o++ = (char)(0x80 | ((cp >> 6) & 0x3F));
        *o++ = (char)(0x80 | (cp & 0x3F));
    } else {
        *o++ = (char)(0xF0 | (cp >> 18));
        *o++ = (char)(0x80 | ((cp >> 12) & 0x3F));
        *o++ = (char)(0x80 | ((cp >> 6) & 0x3F));
        *o++ = (char)(0x80 | (cp & 0x3F));
    }

    *out = o;
}

static char* mj_string(MJParser* ps) {
    /* ps->p at the opening quote */
    const char* s = ++ps->p;
    size_t cap = 16;

    while (*s && *s != '"') {
        if (*s == '\\' && s[1]) {
            s++;
        }

        s++;
        cap++;
    }

    if (*s != '"') {
        mj_fail(ps, "unterminated string");
        return NULL;
    }

    char* buf = (char*)malloc(cap * 4);
    char* o = buf;

    while (*ps->p != '"') {
        char c = *ps->p++;

        if (c != '\\') {
            *o++ = c;
            continue;
        }

        c = *ps->p++;

        switch (c) {
            case 'n':
                *o++ = '\n';
                break;

            case 't':
                *o++ = '\t';
                break;

            case 'r':
                *o++ = '\r';
                break;

            case 'b':
                *o++ = '\b';
                break;

            case 'f':
                *o++ = '\f';
                break;

            case 'u': {
                unsigned cp = 0;

                for (int i = 0; i < 4; i++) {
                    char h = *ps->p++;
                    cp <<= 4;

                    if (h >= '0' && h <= '9') {
                        cp |= (unsigned)(h - '0');
                    } else if (h >= 'a' && h <= 'f') {
                        cp |= (unsigned)(h - 'a' + 10);
                    } else if (h >= 'A' && h <= 'F') {
                        cp |= (unsigned)(h - 'A' + 10);
                    } else {
                        mj_fail(ps, "bad \\u escape");
                        free(buf);
                        return NULL;
                    }
                }

                mj_put_utf8(&o, cp);
                break;
            }

            default:
                *o++ = c;
                break;
        }
    }

    ps->p++; /* closing quote */
    *o = '\0';
    return buf;
}

static MJNode* mj_container(MJParser* ps, int isobj) {
    MJNode* node = mj_new(isobj ? MJ_OBJECT : MJ_ARRAY);
    MJNode* last = NULL;
    char close = isobj ? '}' : ']';
    ps->p++;
    mj_skipws(ps);

    if (*ps->p == close) {
        ps->p++;
        return node;
    }

    while (!ps->failed) {
        char* key = NULL;
        mj_skipws(ps);

        if (isobj) {
            if (*ps->p != '"') {
                mj_fail(ps, "expected a member name");
                break;
            }

            key = mj_string(ps);

            if (!key) {
                break;
            }

            mj_skipws(ps);

            if (*ps->p != ':') {
                free(key);
                mj_fail(ps, "expected ':'");
                break;
            }

            ps->p++;
        }

        MJNode* v = mj_value(ps);

        if (!v) {
            free(key);
            break;
        }

        v->key = key;

        if (last) {
            last->next = v;
        } else {
            node->child = v;
        }

        last = v;
        node->count++;
        mj_skipws(ps);

        if (*ps->p == ',') {
            ps->p++;
            continue;
        }

        if (*ps->p == close) {
            ps->p++;
            return node;
        }

        mj_fail(ps, isobj ? "expected ',' or '}'" : "expected ',' or ']'");
    }

    mj_free(node);
    return NULL;
}

static MJNode* mj_value(MJParser* ps) {
    mj_skipws(ps);
    char c = *ps->p;

    if (c == '{' || c == '[') {
        return mj_container(ps, c == '{');
    }

    if (c == '"') {
        char* s = mj_string(ps);

        if (!s) {
            return NULL;
        }

        MJNode* n = mj_new(MJ_STRING);
        n->str = s;
        return n;
    }

    if (!strncmp(ps->p, "true", 4)) {
        ps->p += 4;
        MJNode* n = mj_new(MJ_TRUE);
        n->num = 1;
        return n;
    }

    if (!strncmp(ps->p, "false", 5)) {
        ps->p += 5;
        return mj_new(MJ_FALSE);
    }

    if (!strncmp(ps->p, "null", 4)) {
        ps->p += 4;
        return mj_new(MJ_NULL);
    }

    if (c == '-' || c == '+' || c == '.' || isdigit((unsigned char)c)) {
        char* end = NULL;
        double v = strtod(ps->p, &end);

        if (end == ps->p) {
            mj_fail(ps, "bad number");
            return NULL;
        }

        ps->p = end;
        MJNode* n = mj_new(MJ_NUMBER);
        n->num = v;
        return n;
    }

    mj_fail(ps, "unexpected character");
    return NULL;
}

MJNode* mj_parse(const char* text, char* err, int errlen) {
    MJParser ps = {text, text, err, errlen, 0};

    if (!text) {
        mj_fail(&ps, "empty input");
        return NULL;
    }

    MJNode* root = mj_value(&ps);

    if (root) {
        mj_skipws(&ps);

        if (*ps.p) {
            mj_fail(&ps, "trailing characters");
            mj_free(root);
            return NULL;
        }
    }

    return root;
}

void mj_free(MJNode* node) {
    while (node) {
        MJNode* next = node->next;
        mj_free(node->child);
        free(node->key);
        free(node->str);
        free(node);
        node = next;
    }
}

MJNode* mj_get(const MJNode* obj, const char* key) {
    if (!obj || obj->type != MJ_OBJECT) {
        return NULL;
    }

    for (MJNode* c = obj->child; c; c = c->next) {
        if (c->key && !strcmp(c->key, key)) {
            return c;
        }
    }

    return NULL;
}

MJNode* mj_at(const MJNode* arr, int idx) {
    if (!arr || (arr->type != MJ_ARRAY && arr->type != MJ_OBJECT)) {
        return NULL;
    }

    MJNode* c = arr->child;

    while (c && idx-- > 0) {
        c = c->next;
    }

    return c;
}

double mj_num(const MJNode* node, double fallback) {
    if (!node) {
        return fallback;
    }

    if (node->type == MJ_NUMBER || node->type == MJ_TRUE || node->type == MJ_FALSE) {
        return node->num;
    }

    return fallback;
}

int mj_int(const MJNode* node, int fallback) {
    if (!node) {
        return fallback;
    }

    if (node->type == MJ_NUMBER || node->type == MJ_TRUE || node->type == MJ_FALSE) {
        /* saturate like cJSON's valueint */
        double v = node->num;

        if (v >= 2147483647.0) {
            return 2147483647;
        }

        if (v <= -2147483648.0) {
            return (-2147483647 - 1);
        }

        return (int)v;
    }

    return fallback;
}

const char* mj_str(const MJNode* node, const char* fallback) {
    if (!node || node->type != MJ_STRING) {
        return fallback;
    }

    return node->str;
}

typedef struct {
    char* buf;
    size_t len, cap;
} MJBuf;

static void mjb_put(MJBuf* b, const char* s, size_t n) {
    if (b->len + n + 1 > b->cap) {
        b->cap = (b->len + n + 1) * 2;
        b->buf = (char*)realloc(b->buf, b->cap);
    }

    memcpy(b->buf + b->len, s, n);
    b->len += n;
    b->buf[b->len] = '\0';
}

static void mjb_str(MJBuf* b, const char* s) {
    mjb_put(b, "\"", 1);

    for (; *s; s++) {
        if (*s == '"' || *s == '\\') {
            mjb_put(b, "\\", 1);
        }

        mjb_put(b, s, 1);
    }

    mjb_put(b, "\"", 1);
}

static void mj_print_rec(const MJNode* n, MJBuf* b) {
    char tmp[64];

    switch (n->type) {
        case MJ_NULL:
            mjb_put(b, "null", 4);
            break;

        case MJ_FALSE:
            mjb_put(b, "false", 5);
            break;

        case MJ_TRUE:
            mjb_put(b, "true", 4);
            break;

        case MJ_NUMBER:
            snprintf(tmp, sizeof(tmp), "%.17g", n->num);
            mjb_put(b, tmp, strlen(tmp));
            break;

        case MJ_STRING:
            mjb_str(b, n->str);
            break;

        default: {
            int isobj = (n->type == MJ_OBJECT);
            mjb_put(b, isobj ? "{" : "[", 1);

            for (const MJNode* c = n->child; c; c = c->next) {
                if (isobj) {
                    mjb_str(b, c->key ? c->key : "");
                    mjb_put(b, ":", 1);
                }

                mj_print_rec(c, b);

                if (c->next) {
                    mjb_put(b, ",", 1);
                }
            }

            mjb_put(b, isobj ? "}" : "]", 1);
        }
    }
}

char* mj_print(const MJNode* node) {
    MJBuf b = {NULL, 0, 0};
    mjb_put(&b, "", 0);

    if (node) {
        mj_print_rec(node, &b);
    }

    return b.buf;
}
