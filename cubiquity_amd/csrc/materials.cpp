// Material metadata: the .toml file that sits next to a .dag (Metadata::load / save,
// src/application/base/metadata.cpp:34-109; Material, metadata.h:11-35) and the viewer's colour
// table built from it (viewer.cpp:49-54).
//
// The reader accepts the part of TOML that the reference writes and edits by hand produce:
// comments, blank lines, bare or quoted keys, basic and literal strings, integers and floats,
// booleans, arrays (also over several lines, with a trailing comma), inline tables, [table] and
// [[array]] headers (no multi-line strings, dates or times). Keys it does not know are skipped. Material tables may carry `name`, `base_color` and,
// as an extension for the path tracer, `emission`.
#include "cqx_internal.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace cqx {
namespace {

struct Value {
    enum Kind { kNone, kString, kNumber, kBool, kArray } kind = kNone;
    std::string str;
    double num = 0.0;
    bool is_int = false;
    std::vector<Value> items;
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}

    // Parses the whole file, calling on_table(name, is_array) for headers and
    // on_key(key, value) for assignments. Returns false with error() set on a syntax error.
    template <typename OnTable, typename OnKey>
    bool run(OnTable on_table, OnKey on_key)
    {
        for (;;) {
            skip_blank_lines();
            if (eof()) return true;
            if (peek() == '[') {
                const bool arr = peek(1) == '[';
                pos_ += arr ? 2 : 1;
                std::string name;
                if (!read_dotted_key(name)) return false;
                skip_ws();
                if (!expect(']') || (arr && !expect(']'))) return false;
                if (!on_table(name, arr)) return false;
            } else {
                std::string key;
                if (!read_dotted_key(key)) return false;
                skip_ws();
                if (!expect('=')) return false;
                skip_ws();
                Value v;
                if (!read_value(v)) return false;
                if (!on_key(key, v)) return false;
            }
            if (!end_of_line()) return false;
        }
    }

    const std::string& error() const { return err_; }
    int line() const
    {
        int l = 1;
        for (size_t i = 0; i < pos_ && i < s_.size(); i++) l += s_[i] == '\n';
        return l;
    }
    bool fail(const char* what)
    {
        if (err_.empty()) err_ = what;
        return false;
    }

private:
    bool eof() const { return pos_ >= s_.size(); }
    char peek(size_t k = 0) const { return pos_ + k < s_.size() ? s_[pos_ + k] : '\0'; }
    void skip_ws()
    {
        while (!eof() && (peek() == ' ' || peek() == '\t')) pos_++;
    }
    void skip_comment()
    {
        if (peek() == '#')
            while (!eof() && peek() != '\n') pos_++;
    }
    void skip_blank_lines()
    {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\r') pos_++;
            if (peek() == '\n') {
                pos_++;
                continue;
            }
            return;
        }
    }
    bool expect(char c)
    {
        if (peek() != c) return fail("unexpected character");
        pos_++;
        return true;
    }
    bool end_of_line()
    {
        skip_ws();
        skip_comment();
        if (peek() == '\r') pos_++;
        if (eof()) return true;
        if (peek() != '\n') return fail("expected the end of the line");
        pos_++;
        return true;
    }
    bool read_string(std::string& out)
    {
        const char q = peek();
        pos_++;
        out.clear();
        while (!eof() && peek() != q) {
            char c = peek();
            if (c == '\n') return fail("unterminated string");
            if (q == '"' && c == '\\') {
                pos_++;
                switch (peek()) {
                case '"': c = '"'; break;
                case '\\': c = '\\'; break;
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case 'r': c = '\r'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                default: return fail("unsupported escape in string");
                }
            }
            out.push_back(c);
            pos_++;
        }
        return expect(q);
    }
    bool read_key_part(std::string& out)
    {
        if (peek() == '"' || peek() == '\'') return read_string(out);
        out.clear();
        while (!eof() && (std::isalnum(static_cast<unsigned char>(peek())) || peek() == '_' || peek() == '-'))
            out.push_back(s_[pos_++]);
        return out.empty() ? fail("expected a key") : true;
    }
    bool read_dotted_key(std::string& out)
    {
        skip_ws();
        std::string part;
        if (!read_key_part(part)) return false;
        out = part;
        for (;;) {
            skip_ws();
            if (peek() != '.') return true;
            pos_++;
            skip_ws();
            if (!read_key_part(part)) return false;
            out += "." + part;
        }
    }
    // Whitespace, newlines and comments inside an array.
    void skip_array_space()
    {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\r' || peek() == '\n') {
                pos_++;
                continue;
            }
            return;
        }
    }
    bool read_value(Value& v)
    {
        const char c = peek();
        if (c == '"' || c == '\'') {
            v.kind = Value::kString;
            return read_string(v.str);
        }
        if (c == '[') {
            pos_++;
            v.kind = Value::kArray;
            for (;;) {
                skip_array_space();
                if (peek() == ']') {
                    pos_++;
                    return true;
                }
                Value item;
                if (!read_value(item)) return false;
                v.items.push_back(item);
                skip_array_space();
                if (peek() == ',') {
                    pos_++;
                    continue;
                }
                if (peek() != ']') return fail("expected ',' or ']' in an array");
            }
        }
        if (c == '{') { // inline table: read and dropped (no material key takes one)
            pos_++;
            skip_ws();
            if (peek() == '}') {
                pos_++;
                return true;
            }
            for (;;) {
                std::string key;
                if (!read_dotted_key(key)) return false;
                skip_ws();
                if (!expect('=')) return false;
                skip_ws();
                Value item;
                if (!read_value(item)) return false;
                skip_ws();
                if (peek() == ',') {
                    pos_++;
                    continue;
                }
                return expect('}');
            }
        }
        if (s_.compare(pos_, 4, "true") == 0 || s_.compare(pos_, 5, "false") == 0) {
            v.kind = Value::kBool;
            v.num = c == 't' ? 1.0 : 0.0;
            pos_ += c == 't' ? 4 : 5;
            return true;
        }
        // number: digits, sign, '.', exponent, '_' separators (removed), inf / nan
        std::string tok;
        while (!eof() && (std::isalnum(static_cast<unsigned char>(peek())) || peek() == '+' || peek() == '-' ||
                          peek() == '.' || peek() == '_'))
            tok.push_back(s_[pos_++]);
        std::string clean;
        for (char ch : tok)
            if (ch != '_') clean.push_back(ch);
        if (clean.empty()) return fail("expected a value");
        v.kind = Value::kNumber;
        if (clean == "inf" || clean == "+inf") v.num = HUGE_VAL;
        else if (clean == "-inf") v.num = -HUGE_VAL;
        else if (clean == "nan" || clean == "+nan" || clean == "-nan") v.num = std::nan("");
        else {
            char* end = nullptr;
            errno = 0;
            v.num = std::strtod(clean.c_str(), &end);
            if (end != clean.c_str() + clean.size() || errno == ERANGE) return fail("malformed number");
            v.is_int = clean.find_first_of(".eEn") == std::string::npos;
        }
        return true;
    }

    const std::string& s_;
    size_t pos_ = 0;
    std::string err_;
};

bool read_vec3(const Value& v, float out[3])
{
    if (v.kind != Value::kArray || v.items.size() != 3) return false;
    for (int i = 0; i < 3; i++) {
        if (v.items[i].kind != Value::kNumber) return false;
        out[i] = float(v.items[i].num);
    }
    return true;
}

} // namespace
} // namespace cqx

using namespace cqx;

extern "C" {

cqx_status cqx_materials_load(const char* path, cqx_material* materials, uint32_t* count, int32_t dimensions[3],
                              uint32_t* has_dimensions)
{
    if (!path || !materials || !count) return fail(CQX_ERR_INVALID_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CQX_ERR_IO, "cannot open %s", path);
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    const bool read_error = std::ferror(f) != 0;
    std::fclose(f);
    if (read_error) return fail(CQX_ERR_IO, "cannot read %s", path);

    uint32_t n = 0;
    bool in_material = false;
    bool dims = false;
    int32_t dim[3] = { 0, 0, 0 };
    Parser p(text);
    const bool ok = p.run(
        [&](const std::string& name, bool is_array) {
            in_material = is_array && name == "materials";
            if (in_material) {
                if (n == CQX_MATERIAL_COUNT) return p.fail("more than 256 materials");
                cqx_material& m = materials[n++];
                std::memset(&m, 0, sizeof(m));
                // Material::DefaultColor, metadata.h:31
                m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.8f;
            }
            return true;
        },
        [&](const std::string& key, const Value& v) {
            if (in_material) {
                cqx_material& m = materials[n - 1];
                if (key == "name") {
                    if (v.kind != Value::kString) return p.fail("material name is not a string");
                    if (v.str.size() >= sizeof(m.name)) return p.fail("material name too long");
                    std::memcpy(m.name, v.str.c_str(), v.str.size() + 1);
                    m.flags |= CQX_MATERIAL_HAS_NAME;
                } else if (key == "base_color") {
                    if (!read_vec3(v, m.base_color)) return p.fail("base_color is not an array of 3 numbers");
                    m.flags |= CQX_MATERIAL_HAS_BASE_COLOR;
                } else if (key == "emission") {
                    if (!read_vec3(v, m.emission)) return p.fail("emission is not an array of 3 numbers");
                    m.flags |= CQX_MATERIAL_HAS_EMISSION;
                }
            } else if (key == "dimensions") {
                if (v.kind != Value::kArray || v.items.size() != 3) return p.fail("dimensions is not an array of 3");
                for (int i = 0; i < 3; i++) {
                    const Value& e = v.items[i];
                    if (e.kind != Value::kNumber || !e.is_int || e.num < INT32_MIN || e.num > INT32_MAX)
                        return p.fail("dimensions are not integers");
                    dim[i] = int32_t(e.num);
                }
                dims = true;
            }
            return true;
        });
    if (!ok) return fail(CQX_ERR_PARSE, "%s:%d: %s", path, p.line(), p.error().c_str());
    *count = n;
    if (dimensions) std::memcpy(dimensions, dim, sizeof(dim));
    if (has_dimensions) *has_dimensions = dims ? 1u : 0u;
    return CQX_OK;
}

cqx_status cqx_materials_save(const char* path, const cqx_material* materials, uint32_t count)
{
    if (!path || (count && !materials)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (count > CQX_MATERIAL_COUNT) return fail(CQX_ERR_LIMIT, "more than 256 materials");
    for (uint32_t i = 0; i < count; i++) {
        const cqx_material& m = materials[i];
        if (std::memchr(m.name, '\0', sizeof(m.name)) == nullptr)
            return fail(CQX_ERR_INVALID_ARG, "material %u: name not terminated", i);
        for (const char* c = m.name; *c; c++)
            if (*c == '"' || *c == '\\' || *c == '\n' || *c == '\r')
                return fail(CQX_ERR_INVALID_ARG, "material %u: name holds a quote, backslash or newline", i);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CQX_ERR_IO, "cannot create %s", path);
    // Metadata::save layout (metadata.cpp:67-109): a comment block, then one [[materials]] table per
    // material with fixed four-decimal colours.
    if (count)
        std::fprintf(f,
                     "# Materials\n# ---------\n# This array of tables contains the material properties. Each voxel "
                     "is an 8-bit\n# unsigned integer and so there can be up to 256 materials in the array "
                     "below.\n\n");
    for (uint32_t i = 0; i < count; i++) {
        const cqx_material& m = materials[i];
        std::fprintf(f, "[[materials]] # Material index %u\n", i);
        if (m.flags & CQX_MATERIAL_HAS_NAME) std::fprintf(f, "name = \"%s\"\n", m.name);
        if (m.flags & CQX_MATERIAL_HAS_BASE_COLOR)
            std::fprintf(f, "base_color = [%.4f, %.4f, %.4f]\n", double(m.base_color[0]), double(m.base_color[1]),
                         double(m.base_color[2]));
        if (m.flags & CQX_MATERIAL_HAS_EMISSION)
            std::fprintf(f, "emission = [%.4f, %.4f, %.4f]\n", double(m.emission[0]), double(m.emission[1]),
                         double(m.emission[2]));
        std::fprintf(f, "\n");
    }
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return fail(CQX_ERR_IO, "cannot write %s", path);
    return CQX_OK;
}

cqx_status cqx_materials_colour_table(const cqx_material* materials, uint32_t count, float colours[256 * 3])
{
    if (!colours || (count && !materials)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (count > CQX_MATERIAL_COUNT) return fail(CQX_ERR_LIMIT, "more than 256 materials");
    for (uint32_t i = 0; i < CQX_MATERIAL_COUNT; i++) {
        const bool defined = i < count;
        for (int c = 0; c < 3; c++)
            colours[i * 3 + c] = defined ? materials[i].base_color[c] : (c == 1 ? 0.0f : 1.0f); // purple
    }
    return CQX_OK;
}

} // extern "C"
