// Graph input formats: graph-tool's binary `.gt` files and delimited edge
// lists, parsed on the host into an edge list in edge-id order (plus property
// columns and vertex names), from which gt_graph_from_doc builds the device
// adjacency with the device builder.
//
// .gt (crates/gt-io/src/gt.rs:1-58, read :555-700):
//   magic "\xe2\x9b\xbe gt" (6 bytes), version u8 == 1, big_endian u8, comment
//   string, directed u8, N u64, then per vertex a u64 count and that many
//   neighbour indices of the narrowest unsigned width holding N (1, 2, 4, 8
//   bytes; :694-705), then nprops u64 and per property: domain u8 (0 graph,
//   1 vertex, 2 edge), name string, value-type index u8, and one value per
//   element of the domain (1, N, or the number of edges).  A string / vector
//   is a u64 count then the elements; every multi-byte field follows the
//   big_endian flag.  Edge i is the i-th adjacency entry of the traversal, so
//   the edge list is (v, u) in file order and edge properties are positional.
// Edge lists (crates/gt-io/src/csv.rs:129-200, dialect :270-420): one record
//   per line, Python's default csv dialect (a field opening with '"' runs to
//   the matching quote, "" is a literal quote, quoted fields may hold the
//   delimiter and newlines), fields stripped of ASCII whitespace, fields past
//   the second parsed and discarded, blank records skipped, an optional
//   header record, a UTF-8 BOM ignored.  Symbolic endpoints are interned in
//   first-seen order, source before target; index endpoints give max + 1
//   vertices.  Record i is edge i.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "gt_hip.h"

#define GT_STATUS_TRY(e)            \
    do {                            \
        gt_status _s = (e);         \
        if (_s != GT_OK) return _s; \
    } while (0)

namespace gt {
gt_status fail(gt_status code, const std::string &msg);
void clear_error();
}  // namespace gt

// Value-type indices of the format (graph-tool's value_types order,
// crates/gt-io/src/gt.rs:150-190).
enum : int32_t {
    kBool = 0, kI16, kI32, kI64, kF64, kLongDouble, kStr, kVecBool, kVecI16, kVecI32,
    kVecI64, kVecF64, kVecLongDouble, kVecStr, kPyObject, kNumKinds
};

struct gt_doc {
    uint64_t n = 0;
    int32_t directed = 1;
    std::string comment;
    std::vector<uint32_t> src, tgt;
    struct Prop {
        std::string name;
        int32_t domain = 0, kind = 0;
        std::vector<double> scalars;             // numeric scalar kinds
        std::vector<std::string> cells;          // every other kind: one blob per element
    };
    std::vector<Prop> props;
    std::vector<std::string> names;              // symbolic edge lists: vertex names
};

namespace {

using gt::fail;

struct Reader {
    const uint8_t *p;
    uint64_t len, pos = 0;
    bool be = false;
    std::string err;

    bool need(uint64_t k) {
        if (len - pos >= k) return true;
        err = "at byte " + std::to_string(pos) + ": unexpected end of file, wanted " +
              std::to_string(k) + " more bytes";
        return false;
    }
    template <typename T>
    bool scalar(T *out) {
        if (!need(sizeof(T))) return false;
        uint8_t b[sizeof(T)];
        std::memcpy(b, p + pos, sizeof(T));
        pos += sizeof(T);
        const bool host_be = false;  // gfx950 hosts are little-endian x86-64
        if (be != host_be)
            for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
        std::memcpy(out, b, sizeof(T));
        return true;
    }
    bool bytes(std::string *out) {
        uint64_t k = 0;
        if (!scalar(&k)) return false;
        if (len - pos < k) {
            err = "at byte " + std::to_string(pos) + ": string declares " + std::to_string(k) +
                  " bytes, " + std::to_string(len - pos) + " are present";
            return false;
        }
        out->assign(reinterpret_cast<const char *>(p + pos), (size_t)k);
        pos += k;
        return true;
    }
    bool count(uint64_t *k) {
        if (!scalar(k)) return false;
        return true;
    }
};

int vint_width(uint64_t n) {
    if (n <= 0xffull) return 1;
    if (n <= 0xffffull) return 2;
    if (n <= 0xffffffffull) return 4;
    return 8;
}

bool read_vint(Reader &r, int w, uint64_t *v) {
    switch (w) {
        case 1: { uint8_t x; if (!r.scalar(&x)) return false; *v = x; return true; }
        case 2: { uint16_t x; if (!r.scalar(&x)) return false; *v = x; return true; }
        case 4: { uint32_t x; if (!r.scalar(&x)) return false; *v = x; return true; }
        default: return r.scalar(v);
    }
}

template <typename T>
void put(std::string &blob, T v) {
    blob.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

// One element of a column.  Numeric scalars go to `num`; everything else to a
// blob: strings and pickles as their bytes, long doubles as their 16 bytes,
// numeric vectors as packed f64, vectors of long double as packed 16-byte
// values, vectors of strings as (u64 length, bytes) runs after a u64 count.
bool read_cell(Reader &r, int32_t kind, double *num, std::string *blob) {
    switch (kind) {
        case kBool: { uint8_t x; if (!r.scalar(&x)) return false; *num = x; return true; }
        case kI16: { int16_t x; if (!r.scalar(&x)) return false; *num = x; return true; }
        case kI32: { int32_t x; if (!r.scalar(&x)) return false; *num = x; return true; }
        case kI64: { int64_t x; if (!r.scalar(&x)) return false; *num = (double)x; return true; }
        case kF64: return r.scalar(num);
        case kLongDouble: {
            if (!r.need(16)) return false;
            uint8_t b[16];
            std::memcpy(b, r.p + r.pos, 16);
            r.pos += 16;
            if (r.be) std::reverse(b, b + 16);
            blob->assign(reinterpret_cast<const char *>(b), 16);
            return true;
        }
        case kStr:
        case kPyObject:
        case kVecBool: {
            if (!r.bytes(blob)) return false;
            if (kind == kVecBool) {  // widen to packed f64 like the other vectors
                std::string d;
                for (unsigned char c : *blob) put<double>(d, (double)c);
                *blob = d;
            }
            return true;
        }
        case kVecI16: case kVecI32: case kVecI64: case kVecF64: {
            uint64_t k;
            if (!r.count(&k)) return false;
            blob->clear();
            for (uint64_t i = 0; i < k; ++i) {
                double v = 0;
                if (kind == kVecI16) { int16_t x; if (!r.scalar(&x)) return false; v = x; }
                else if (kind == kVecI32) { int32_t x; if (!r.scalar(&x)) return false; v = x; }
                else if (kind == kVecI64) { int64_t x; if (!r.scalar(&x)) return false; v = (double)x; }
                else if (!r.scalar(&v)) return false;
                put<double>(*blob, v);
            }
            return true;
        }
        case kVecLongDouble: {
            uint64_t k;
            if (!r.count(&k)) return false;
            blob->clear();
            for (uint64_t i = 0; i < k; ++i) {
                if (!r.need(16)) return false;
                uint8_t b[16];
                std::memcpy(b, r.p + r.pos, 16);
                r.pos += 16;
                if (r.be) std::reverse(b, b + 16);
                blob->append(reinterpret_cast<const char *>(b), 16);
            }
            return true;
        }
        case kVecStr: {
            uint64_t k;
            if (!r.count(&k)) return false;
            blob->clear();
            put<uint64_t>(*blob, k);
            for (uint64_t i = 0; i < k; ++i) {
                std::string s;
                if (!r.bytes(&s)) return false;
                put<uint64_t>(*blob, s.size());
                blob->append(s);
            }
            return true;
        }
        default: return false;
    }
}

bool numeric_scalar(int32_t kind) { return kind <= kF64 && kind != kLongDouble; }

gt_status parse_gt(const uint8_t *p, uint64_t len, gt_doc *d) {
    static const uint8_t magic[6] = {0xe2, 0x9b, 0xbe, 0x20, 0x67, 0x74};
    if (len < 6 || std::memcmp(p, magic, 6) != 0)
        return fail(GT_ERR_INVALID_ARG, "not a .gt file: bad magic");
    Reader r{p, len};
    r.pos = 6;
    uint8_t version = 0, be = 0;
    if (!r.scalar(&version)) return fail(GT_ERR_INVALID_ARG, r.err);
    if (version != 1)
        return fail(GT_ERR_INVALID_ARG, "unsupported .gt version " + std::to_string(version));
    if (!r.scalar(&be)) return fail(GT_ERR_INVALID_ARG, r.err);
    r.be = be != 0;
    if (!r.bytes(&d->comment)) return fail(GT_ERR_INVALID_ARG, r.err);
    uint8_t dir = 0;
    uint64_t n = 0;
    if (!r.scalar(&dir) || !r.scalar(&n)) return fail(GT_ERR_INVALID_ARG, r.err);
    if (n > 0xffffffffull)
        return fail(GT_ERR_INVALID_ARG, "vertex count exceeds the 32-bit index width");
    d->directed = dir != 0;
    d->n = n;
    const int w = vint_width(n);
    for (uint64_t v = 0; v < n; ++v) {
        uint64_t k = 0;
        if (!r.count(&k)) return fail(GT_ERR_INVALID_ARG, r.err);
        if (k > (len - r.pos) / (uint64_t)w)
            return fail(GT_ERR_INVALID_ARG, "at byte " + std::to_string(r.pos) + ": vertex " +
                                                std::to_string(v) + " declares " +
                                                std::to_string(k) + " neighbours, the file ends first");
        for (uint64_t i = 0; i < k; ++i) {
            uint64_t u = 0;
            if (!read_vint(r, w, &u)) return fail(GT_ERR_INVALID_ARG, r.err);
            if (u >= n)
                return fail(GT_ERR_INVALID_ARG, "vertex index " + std::to_string(u) +
                                                    " out of range (bound " + std::to_string(n) +
                                                    ")");
            d->src.push_back((uint32_t)v);
            d->tgt.push_back((uint32_t)u);
        }
        if (d->src.size() > 0xffffffffull)
            return fail(GT_ERR_INVALID_ARG, "edge count exceeds the 32-bit index width");
    }
    uint64_t np = 0;
    if (!r.scalar(&np)) return fail(GT_ERR_INVALID_ARG, r.err);
    for (uint64_t k = 0; k < np; ++k) {
        gt_doc::Prop pr;
        uint8_t dom = 0, kind = 0;
        const uint64_t at = r.pos;
        if (!r.scalar(&dom)) return fail(GT_ERR_INVALID_ARG, r.err);
        if (dom > 2)
            return fail(GT_ERR_INVALID_ARG, "at byte " + std::to_string(at) +
                                                ": invalid property domain " + std::to_string(dom));
        if (!r.bytes(&pr.name)) return fail(GT_ERR_INVALID_ARG, r.err);
        if (!r.scalar(&kind)) return fail(GT_ERR_INVALID_ARG, r.err);
        if (kind >= kNumKinds)
            return fail(GT_ERR_INVALID_ARG, "unknown value type index " + std::to_string(kind));
        pr.domain = dom;
        pr.kind = kind;
        const uint64_t cnt = dom == 0 ? 1 : dom == 1 ? n : (uint64_t)d->src.size();
        for (uint64_t i = 0; i < cnt; ++i) {
            double num = 0;
            std::string blob;
            if (!read_cell(r, kind, &num, &blob)) return fail(GT_ERR_INVALID_ARG, r.err);
            if (numeric_scalar(kind)) pr.scalars.push_back(num);
            else pr.cells.push_back(std::move(blob));
        }
        d->props.push_back(std::move(pr));
    }
    return GT_OK;
}

// ---------------------------------------------------------------- edge lists

enum class Stop { Field, Record, Eof };

struct Csv {
    const uint8_t *b;
    uint64_t len, pos = 0, line = 1;
    uint8_t delim;

    void eat_terminator() {
        if (pos < len && b[pos] == '\r') {
            ++pos;
            if (pos < len && b[pos] == '\n') ++pos;
            ++line;
        } else if (pos < len && b[pos] == '\n') {
            ++pos;
            ++line;
        }
    }
    // Length of the White_Space code point starting at s[i] (0 if none): the
    // set Rust's str::trim removes -- ASCII tab..CR and space, U+0085, U+00A0,
    // U+1680, U+2000-200A, U+2028, U+2029, U+202F, U+205F, U+3000.
    static size_t ws_at(const std::string &s, size_t i) {
        const unsigned char c = (unsigned char)s[i];
        if (c == ' ' || (c >= 9 && c <= 13)) return 1;
        auto at = [&](size_t k) { return i + k < s.size() ? (unsigned char)s[i + k] : 0u; };
        if (c == 0xc2 && (at(1) == 0x85 || at(1) == 0xa0)) return 2;
        if (c == 0xe1 && at(1) == 0x9a && at(2) == 0x80) return 3;
        if (c == 0xe2 && at(1) == 0x80 &&
            (at(2) <= 0x8a || at(2) == 0xa8 || at(2) == 0xa9 || at(2) == 0xaf))
            return 3;
        if (c == 0xe2 && at(1) == 0x81 && at(2) == 0x9f) return 3;
        if (c == 0xe3 && at(1) == 0x80 && at(2) == 0x80) return 3;
        return 0;
    }
    static std::string strip(const std::string &s) {
        size_t a = 0, e = s.size();
        for (size_t k; a < e && (k = ws_at(s, a)) != 0;) a += k;
        // From the right: step back over one code point at a time.
        while (e > a) {
            size_t b = e - 1;
            while (b > a && (((unsigned char)s[b]) & 0xc0) == 0x80) --b;
            const size_t k = ws_at(s, b);
            if (k == 0 || b + k != e) break;
            e = b;
        }
        return s.substr(a, e - a);
    }
    // One field; the cursor ends after its delimiter or on the terminator.
    gt_status field(std::string *out, Stop *stop) {
        const uint64_t start = pos, start_line = line;
        bool quoted = false, in_q = false;
        std::string acc;
        for (;;) {
            if (pos >= len) {
                *stop = Stop::Eof;
                break;
            }
            const uint8_t c = b[pos];
            if (in_q) {
                if (c == '"') {
                    if (pos + 1 < len && b[pos + 1] == '"') {
                        acc.push_back('"');
                        pos += 2;
                    } else {
                        in_q = false;
                        ++pos;
                    }
                } else {
                    if (c == '\n') ++line;
                    acc.push_back((char)c);
                    ++pos;
                }
            } else if (c == delim) {
                ++pos;
                *stop = Stop::Field;
                break;
            } else if (c == '\n' || c == '\r') {
                *stop = Stop::Record;
                break;
            } else if (c == '"' && pos == start) {
                quoted = in_q = true;
                ++pos;
            } else {
                acc.push_back((char)c);
                ++pos;
            }
        }
        if (in_q) return fail(GT_ERR_INVALID_ARG, "line " + std::to_string(start_line) +
                                                      ": unterminated quoted field");
        (void)quoted;
        *out = strip(acc);
        return GT_OK;
    }
    gt_status drain() {
        for (;;) {
            std::string f;
            Stop st;
            GT_STATUS_TRY(field(&f, &st));
            if (st == Stop::Field) continue;
            if (st == Stop::Record) eat_terminator();
            return GT_OK;
        }
    }
};

bool valid_utf8(const std::string &s) {
    const size_t n = s.size();
    size_t i = 0;
    while (i < n) {
        const unsigned char c = (unsigned char)s[i];
        size_t k;
        if (c < 0x80) k = 0;
        else if ((c >> 5) == 0x6) k = 1;
        else if ((c >> 4) == 0xe) k = 2;
        else if ((c >> 3) == 0x1e) k = 3;
        else return false;
        if (i + k >= n && k > 0) return false;
        for (size_t j = 1; j <= k; ++j)
            if ((((unsigned char)s[i + j]) & 0xc0) != 0x80) return false;
        i += k + 1;
    }
    return true;
}

gt_status parse_edge_list(const uint8_t *p, uint64_t len, uint8_t delim, bool header,
                          bool symbolic, gt_doc *d) {
    if (len >= 3 && p[0] == 0xef && p[1] == 0xbb && p[2] == 0xbf) {  // UTF-8 BOM
        p += 3;
        len -= 3;
    }
    Csv c{p, len, 0, 1, delim};
    if (header) {
        while (c.pos < c.len && (c.b[c.pos] == '\n' || c.b[c.pos] == '\r')) c.eat_terminator();
        if (c.pos < c.len) GT_STATUS_TRY(c.drain());
    }
    std::unordered_map<std::string, uint32_t> index;
    uint64_t nv = 0;
    for (;;) {
        if (c.pos >= c.len) break;
        if (c.b[c.pos] == '\n' || c.b[c.pos] == '\r') {
            c.eat_terminator();
            continue;
        }
        const uint64_t line = c.line;
        std::string s, t;
        Stop st;
        GT_STATUS_TRY(c.field(&s, &st));
        if (st != Stop::Field) {
            if (st == Stop::Record) c.eat_terminator();
            if (s.empty()) continue;  // a line of whitespace is blank
            return fail(GT_ERR_INVALID_ARG,
                        "line " + std::to_string(line) + ": a record needs at least two fields");
        }
        GT_STATUS_TRY(c.field(&t, &st));
        if (st == Stop::Field) GT_STATUS_TRY(c.drain());
        else if (st == Stop::Record) c.eat_terminator();
        uint32_t ids[2];
        const std::string *f[2] = {&s, &t};
        for (int k = 0; k < 2; ++k) {
            if (!valid_utf8(*f[k]))
                return fail(GT_ERR_INVALID_ARG,
                            "line " + std::to_string(line) + ": field is not valid UTF-8");
            if (symbolic) {
                auto it = index.find(*f[k]);
                if (it == index.end()) {
                    if (d->names.size() >= 0xffffffffull)
                        return fail(GT_ERR_INVALID_ARG, "vertex count exceeds the index width");
                    it = index.emplace(*f[k], (uint32_t)d->names.size()).first;
                    d->names.push_back(*f[k]);
                }
                ids[k] = it->second;
            } else {
                const std::string &x0 = *f[k];
                // usize parsing accepts one leading '+'.
                const std::string x = (!x0.empty() && x0[0] == '+') ? x0.substr(1) : x0;
                bool ok = !x.empty();
                uint64_t v = 0;
                for (char ch : x) {
                    if (ch < '0' || ch > '9' || v > 0x0fffffffffffffffull) {
                        ok = false;
                        break;
                    }
                    v = v * 10 + (uint64_t)(ch - '0');
                }
                if (!ok)
                    return fail(GT_ERR_INVALID_ARG, "line " + std::to_string(line) +
                                                        ": expected a vertex index, found \"" +
                                                        x0 + "\"");
                if (v >= 0xffffffffull)
                    return fail(GT_ERR_INVALID_ARG, "vertex index exceeds the index width");
                ids[k] = (uint32_t)v;
                nv = std::max<uint64_t>(nv, v + 1);
            }
        }
        d->src.push_back(ids[0]);
        d->tgt.push_back(ids[1]);
        if (d->src.size() > 0xffffffffull)
            return fail(GT_ERR_INVALID_ARG, "edge count exceeds the index width");
    }
    d->n = symbolic ? d->names.size() : nv;
    d->directed = 1;
    return GT_OK;
}

}  // namespace

extern "C" {

gt_status gt_read_gt(const void *bytes, uint64_t len, gt_doc **out) {
    gt::clear_error();
    if (!out || (!bytes && len)) return fail(GT_ERR_INVALID_ARG, "null argument");
    gt_doc *d = new gt_doc();
    const gt_status st = parse_gt(static_cast<const uint8_t *>(bytes), len, d);
    if (st != GT_OK) {
        delete d;
        return st;
    }
    *out = d;
    return GT_OK;
}

gt_status gt_read_edge_list(const void *bytes, uint64_t len, uint8_t delimiter, int32_t header,
                            int32_t symbolic, gt_doc **out) {
    gt::clear_error();
    if (!out || (!bytes && len)) return fail(GT_ERR_INVALID_ARG, "null argument");
    if (delimiter == '"' || delimiter == '\n' || delimiter == '\r')
        return fail(GT_ERR_INVALID_ARG, "the delimiter cannot be a quote or a line break");
    gt_doc *d = new gt_doc();
    const gt_status st = parse_edge_list(static_cast<const uint8_t *>(bytes), len, delimiter,
                                         header != 0, symbolic != 0, d);
    if (st != GT_OK) {
        delete d;
        return st;
    }
    *out = d;
    return GT_OK;
}

gt_status gt_doc_free(gt_doc *d) {
    delete d;
    return GT_OK;
}

gt_status gt_doc_info(const gt_doc *d, uint64_t *n, uint64_t *m, int32_t *directed,
                      uint32_t *nprops, uint64_t *nnames) {
    gt::clear_error();
    if (!d) return fail(GT_ERR_INVALID_ARG, "null document");
    if (n) *n = d->n;
    if (m) *m = d->src.size();
    if (directed) *directed = d->directed;
    if (nprops) *nprops = (uint32_t)d->props.size();
    if (nnames) *nnames = d->names.size();
    return GT_OK;
}

gt_status gt_doc_edges(const gt_doc *d, uint32_t *src, uint32_t *tgt) {
    gt::clear_error();
    if (!d) return fail(GT_ERR_INVALID_ARG, "null document");
    if (!d->src.empty() && (!src || !tgt)) return fail(GT_ERR_INVALID_ARG, "null output");
    if (!d->src.empty()) {
        std::memcpy(src, d->src.data(), d->src.size() * sizeof(uint32_t));
        std::memcpy(tgt, d->tgt.data(), d->tgt.size() * sizeof(uint32_t));
    }
    return GT_OK;
}

gt_status gt_doc_comment(const gt_doc *d, const char **s, uint64_t *len) {
    gt::clear_error();
    if (!d || !s || !len) return fail(GT_ERR_INVALID_ARG, "null argument");
    *s = d->comment.data();
    *len = d->comment.size();
    return GT_OK;
}

gt_status gt_doc_prop_info(const gt_doc *d, uint32_t k, const char **name, uint64_t *name_len,
                           int32_t *domain, int32_t *kind, uint64_t *count) {
    gt::clear_error();
    if (!d) return fail(GT_ERR_INVALID_ARG, "null document");
    if (k >= d->props.size()) return fail(GT_ERR_INVALID_ARG, "no such property");
    const gt_doc::Prop &p = d->props[k];
    if (name) *name = p.name.data();
    if (name_len) *name_len = p.name.size();
    if (domain) *domain = p.domain;
    if (kind) *kind = p.kind;
    if (count) *count = numeric_scalar(p.kind) ? p.scalars.size() : p.cells.size();
    return GT_OK;
}

gt_status gt_doc_prop_f64(const gt_doc *d, uint32_t k, double *out) {
    gt::clear_error();
    if (!d) return fail(GT_ERR_INVALID_ARG, "null document");
    if (k >= d->props.size()) return fail(GT_ERR_INVALID_ARG, "no such property");
    const gt_doc::Prop &p = d->props[k];
    if (!numeric_scalar(p.kind))
        return fail(GT_ERR_INVALID_ARG, "property " + p.name + " is not a numeric scalar column");
    if (!p.scalars.empty()) {
        if (!out) return fail(GT_ERR_INVALID_ARG, "null output");
        std::memcpy(out, p.scalars.data(), p.scalars.size() * sizeof(double));
    }
    return GT_OK;
}

gt_status gt_doc_prop_cell(const gt_doc *d, uint32_t k, uint64_t i, const void **data,
                           uint64_t *len) {
    gt::clear_error();
    if (!d || !data || !len) return fail(GT_ERR_INVALID_ARG, "null argument");
    if (k >= d->props.size()) return fail(GT_ERR_INVALID_ARG, "no such property");
    const gt_doc::Prop &p = d->props[k];
    if (numeric_scalar(p.kind))
        return fail(GT_ERR_INVALID_ARG, "property " + p.name + " is a numeric scalar column");
    if (i >= p.cells.size()) return fail(GT_ERR_INVALID_ARG, "element out of range");
    *data = p.cells[i].data();
    *len = p.cells[i].size();
    return GT_OK;
}

gt_status gt_graph_from_doc(const gt_doc *d, uint64_t graph_id, gt_graph **out) {
    gt::clear_error();
    if (!d || !out) return fail(GT_ERR_INVALID_ARG, "null argument");
    return gt_graph_build_edges(graph_id, (uint32_t)d->n, d->src.size(), d->src.data(),
                                d->tgt.data(), GT_HOST, out);
}

gt_status gt_doc_vertex_name(const gt_doc *d, uint64_t v, const char **s, uint64_t *len) {
    gt::clear_error();
    if (!d || !s || !len) return fail(GT_ERR_INVALID_ARG, "null argument");
    if (v >= d->names.size()) return fail(GT_ERR_INVALID_ARG, "no such vertex name");
    *s = d->names[v].data();
    *len = d->names[v].size();
    return GT_OK;
}

}  // extern "C"
