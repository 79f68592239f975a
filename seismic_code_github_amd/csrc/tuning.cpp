// Autotuner cache of a context as a JSON file (no counterpart in the
// reference, whose launch shapes are the OpenMP loops of the CPU code).
//
// Format:
//   {"seisfd_tuning": 1,
//    "entries": [{"arch": "gfx950:sramecc+:xnack-", "class": "acoustic:R8:cpml:x12:z11",
//                 "kind": 0, "a": 16, "b": 4, "us_per_step": 60.7}, ...]}
// Entries measured on another architecture are kept in the file but not used.
// Saving merges the context's entries into an existing file.
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "common.h"

namespace sfd {
namespace {

// Minimal JSON reader: objects, arrays, strings (\" and \\ escapes), numbers,
// true / false / null.  Values are kept as a tagged tree.
struct JValue {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;
    const JValue* get(const char* key) const {
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;

    void ws() {
        while (p < end && std::isspace((unsigned char)*p)) ++p;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(end - p) < n || std::strncmp(p, w, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string& out) {
        if (p >= end || *p != '"') return fail("expected a string");
        ++p;
        out.clear();
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) return fail("unterminated escape");
                if (*p != '"' && *p != '\\' && *p != '/') return fail("unsupported escape");
            }
            out += *p++;
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(JValue& v, int depth) {
        if (depth > 16) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        if (*p == '{') {
            ++p;
            v.kind = JValue::OBJ;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                v.obj.emplace_back(key, JValue());
                if (!value(v.obj.back().second, depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = JValue::ARR;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = JValue::STR;
            return string(v.str);
        }
        if (lit("true")) { v.kind = JValue::BOOL; v.num = 1; return true; }
        if (lit("false")) { v.kind = JValue::BOOL; v.num = 0; return true; }
        if (lit("null")) { v.kind = JValue::NUL; return true; }
        std::string num;
        while (p < end && (std::isdigit((unsigned char)*p) || std::strchr("+-.eE", *p))) num += *p++;
        if (num.empty()) return fail("unexpected character");
        char* e = nullptr;
        v.kind = JValue::NUM;
        v.num = std::strtod(num.c_str(), &e);
        if (!e || *e) return fail("bad number");
        return true;
    }
};

// Text as a JSON string body (quotes and backslashes escaped, control
// characters dropped).
std::string esc(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c >= 0x20) o += c;
    }
    return o;
}

bool as_int(const JValue* v, int* out) {
    if (!v || v->kind != JValue::NUM || v->num != (double)(int)v->num) return false;
    *out = (int)v->num;
    return true;
}

}  // namespace
}  // namespace sfd

using namespace sfd;

extern "C" {

int sfd_context_save_tuning(sfd_context* ctx, const char* path) {
    if (!ctx || !path) return SFD_EINVAL;
    // Entries already in the file are kept unless the context has one for
    // the same architecture and class.
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        if (in) {
            std::stringstream ss;
            ss << in.rdbuf();
            const std::string text = ss.str();
            JParser jp{text.data(), text.data() + text.size(), {}};
            JValue root;
            if (jp.value(root, 0) && root.kind == JValue::OBJ) {
                const JValue* ent = root.get("entries");
                if (ent && ent->kind == JValue::ARR)
                    for (auto& e : ent->arr) {
                        const JValue* a = e.get("arch");
                        const JValue* c = e.get("class");
                        int kind, aa, bb;
                        if (!a || a->kind != JValue::STR || !c || c->kind != JValue::STR ||
                            (a->str == ctx->arch && ctx->tuning.count(c->str)) ||
                            !as_int(e.get("kind"), &kind) || !as_int(e.get("a"), &aa) || !as_int(e.get("b"), &bb))
                            continue;
                        const JValue* us = e.get("us_per_step");
                        char buf[512];
                        snprintf(buf, sizeof buf,
                                 "  {\"arch\": \"%s\", \"class\": \"%s\", \"kind\": %d, \"a\": %d, \"b\": %d, "
                                 "\"us_per_step\": %.3f}",
                                 esc(a->str).c_str(), esc(c->str).c_str(), kind, aa, bb,
                                 (us && us->kind == JValue::NUM) ? us->num : 0.0);
                        keep.push_back(buf);
                    }
            }
        }
    }
    for (auto& [cls, t] : ctx->tuning) {
        char buf[512];
        snprintf(buf, sizeof buf,
                 "  {\"arch\": \"%s\", \"class\": \"%s\", \"kind\": %d, \"a\": %d, \"b\": %d, \"us_per_step\": %.3f}",
                 esc(ctx->arch).c_str(), esc(cls).c_str(), t.kind, t.a, t.b, (double)t.us);
        keep.push_back(buf);
    }
    std::ofstream out(path, std::ios::trunc);
    if (!out) return fail(ctx, SFD_EINVAL, "cannot write %s", path);
    out << "{\"seisfd_tuning\": 1,\n \"entries\": [\n";
    for (size_t i = 0; i < keep.size(); ++i) out << keep[i] << (i + 1 < keep.size() ? ",\n" : "\n");
    out << " ]}\n";
    if (!out) return fail(ctx, SFD_EINVAL, "cannot write %s", path);
    return SFD_OK;
}

int sfd_context_load_tuning(sfd_context* ctx, const char* path) {
    if (!ctx || !path) return SFD_EINVAL;
    std::ifstream in(path);
    if (!in) return fail(ctx, SFD_EINVAL, "cannot read %s", path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    JParser jp{text.data(), text.data() + text.size(), {}};
    JValue root;
    if (!jp.value(root, 0)) return fail(ctx, SFD_EINVAL, "%s: not JSON (%s)", path, jp.err.c_str());
    jp.ws();
    if (jp.p != jp.end) return fail(ctx, SFD_EINVAL, "%s: trailing characters after the JSON value", path);
    int version = 0;
    if (root.kind != JValue::OBJ || !as_int(root.get("seisfd_tuning"), &version) || version != 1)
        return fail(ctx, SFD_EINVAL, "%s: not a seisfd tuning file (version 1)", path);
    const JValue* ent = root.get("entries");
    if (!ent || ent->kind != JValue::ARR) return fail(ctx, SFD_EINVAL, "%s: no entries array", path);
    std::map<std::string, TunedShape> got;
    for (size_t i = 0; i < ent->arr.size(); ++i) {
        const JValue& e = ent->arr[i];
        const JValue* a = e.get("arch");
        const JValue* c = e.get("class");
        TunedShape t;
        if (e.kind != JValue::OBJ || !a || a->kind != JValue::STR || !c || c->kind != JValue::STR ||
            !as_int(e.get("kind"), &t.kind) || !as_int(e.get("a"), &t.a) || !as_int(e.get("b"), &t.b) ||
            t.kind < 0 || t.kind > 2 || t.a <= 0 || t.b <= 0)
            return fail(ctx, SFD_EINVAL, "%s: entry %zu is malformed", path, i);
        const JValue* us = e.get("us_per_step");
        t.us = (us && us->kind == JValue::NUM) ? (float)us->num : 0.f;
        if (a->str == ctx->arch) got[c->str] = t;
    }
    // Entries replace the context's for the same class; a malformed file
    // changes nothing.
    for (auto& kv : got) ctx->tuning[kv.first] = kv.second;
    return SFD_OK;
}

int sfd_context_tuning_size(const sfd_context* ctx) { return ctx ? (int)ctx->tuning.size() : 0; }

}  // extern "C"
