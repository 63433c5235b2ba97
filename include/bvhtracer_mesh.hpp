/*
 * bvhtracer_mesh.hpp -- mesh ingestion for the C++ host API: the
 * reference's `.tri` loader (bvhtracer/tri_loader/src/lexer.rs, loader.rs)
 * and its TriMeshDecoder / ObjMeshDecoder (bvhtracer/src/mesh/decoders.rs:
 * 88-215), producing bvhtracer::Mesh.  Host code only (text parsing); the
 * meshes then go to the GPU through ModelBuilder / BvhBuilder.
 *
 * Numbers convert with strtof, which rounds correctly like Rust's
 * str::parse::<f32>; tokens Rust would reject (hexadecimal, digit
 * separators, trailing characters) are rejected first.
 */
#ifndef BVHTRACER_MESH_HPP
#define BVHTRACER_MESH_HPP

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <istream>
#include <iterator>
#include <sstream>
#include <string>
#include <vector>

#include "bvhtracer.hpp"

namespace bvhtracer {

/* tri_loader's TriLoaderError (loader.rs:40-70). */
class TriLoaderError : public Error {
public:
    enum Kind { ExpectedFloat, EndOfFile };
    TriLoaderError(size_t line, Kind kind, const std::string& message)
        : Error(BVHT_EINVAL, "Tri file format parsing error at line " + std::to_string(line) + ": " + message),
          line_(line), kind_(kind)
    {
    }
    size_t line_number() const { return line_; }
    Kind kind() const { return kind_; }

private:
    size_t line_;
    Kind kind_;
};

namespace detail {

inline bool lower_equals(const std::string& s, size_t from, const char* word)
{
    size_t n = std::char_traits<char>::length(word);
    if (s.size() - from != n)
        return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[from + i]) != word[i])
            return false;
    return true;
}

/* A token Rust's f32 parse accepts: [+-] then inf / infinity / nan (any
 * case), or decimal digits with at most one '.', at least one digit, and an
 * optional exponent [eE][+-]digits. */
inline bool is_number(const std::string& t)
{
    size_t i = 0;
    if (i < t.size() && (t[i] == '+' || t[i] == '-'))
        ++i;
    if (lower_equals(t, i, "inf") || lower_equals(t, i, "infinity") || lower_equals(t, i, "nan"))
        return true;
    size_t digits = 0;
    bool dot = false;
    for (; i < t.size(); ++i) {
        if (std::isdigit((unsigned char)t[i]))
            ++digits;
        else if (t[i] == '.' && !dot)
            dot = true;
        else
            break;
    }
    if (digits == 0)
        return false;
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
        ++i;
        if (i < t.size() && (t[i] == '+' || t[i] == '-'))
            ++i;
        size_t exp_digits = 0;
        while (i < t.size() && std::isdigit((unsigned char)t[i]))
            ++i, ++exp_digits;
        if (exp_digits == 0)
            return false;
    }
    return i == t.size();
}

inline float to_f32(const std::string& t)
{
    errno = 0;
    return std::strtof(t.c_str(), nullptr);  // correctly rounded; overflow gives +-inf like Rust
}

}  // namespace detail

/* tri_loader's lexer and TriLoader::load (lexer.rs:35-104, loader.rs:
 * 108-186): whitespace is ' ', '\\' and '\t'; each '\n' or '\r' is a token
 * of its own; '#' comments to the end of the line; triangles are nine
 * numbers separated by '\n' tokens. */
inline std::vector<Triangle> load_tri_text(const std::string& text)
{
    struct Tok {
        std::string s;
    };
    std::vector<Tok> toks;
    const std::string ws = " \\\t", nl = "\n\r";
    size_t i = 0, n = text.size();
    while (i < n) {
        while (i < n && ws.find(text[i]) != std::string::npos)
            ++i;
        if (i < n && text[i] == '#')
            while (i < n && nl.find(text[i]) == std::string::npos)
                ++i;
        if (i >= n)
            break;
        if (nl.find(text[i]) != std::string::npos) {
            toks.push_back({std::string(1, text[i])});
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && ws.find(text[j]) == std::string::npos && nl.find(text[j]) == std::string::npos)
            ++j;
        toks.push_back({text.substr(i, j - i)});
        i = j;
    }
    std::vector<Triangle> out;
    size_t k = 0, line = 1;
    while (k < toks.size()) {
        if (toks[k].s == "\n") {
            ++line, ++k;
            continue;
        }
        float v[9];
        for (int c = 0; c < 9; ++c) {
            if (k >= toks.size())
                throw TriLoaderError(line, TriLoaderError::EndOfFile,
                                     "Reached the end of the input in the process of getting the next token.");
            const std::string& t = toks[k++].s;
            if (t == "\n")
                ++line;
            if (!detail::is_number(t))
                throw TriLoaderError(line, TriLoaderError::ExpectedFloat,
                                     "Expected a floating point number but got `" + t + "` instead.");
            v[c] = detail::to_f32(t);
        }
        out.emplace_back(Vector3(v[0], v[1], v[2]), Vector3(v[3], v[4], v[5]), Vector3(v[6], v[7], v[8]));
    }
    return out;
}

/* TriMeshDecoder (decoders.rs:88-133): triangles from a .tri document, zero
 * texture coordinates, and at each corner the face normal
 * normalize(normalize(v2 - v0) x normalize(v1 - v0)). */
class TriMeshDecoder {
public:
    explicit TriMeshDecoder(std::istream& reader) : reader_(reader) {}
    Mesh read_mesh()
    {
        std::string text((std::istreambuf_iterator<char>(reader_)), std::istreambuf_iterator<char>());
        MeshBuilder builder;
        for (const Triangle& t : load_tri_text(text)) {
            Vector3 nrm = (t.vertices[2] - t.vertices[0]).normalize().cross((t.vertices[1] - t.vertices[0]).normalize())
                              .normalize();
            Normals normals;
            normals.normals = {nrm, nrm, nrm};
            builder.with_primitive(t, TextureCoordinates{}, normals);
        }
        return builder.build();
    }

private:
    std::istream& reader_;
};

/* ObjMeshDecoder (decoders.rs:136-215) over the OBJ subset the reference
 * reads: v, vt, vn and f records with v, v/vt, v//vn or v/vt/vn corners
 * (1-based or negative relative indices), faces of the first object only,
 * fan-triangulated; missing texture coordinates and normals are zero; '\\'
 * counts as whitespace; other records are skipped. */
class ObjMeshDecoder {
public:
    explicit ObjMeshDecoder(std::istream& reader) : reader_(reader) {}
    Mesh read_mesh()
    {
        std::vector<Vector3> v, vn;
        std::vector<std::array<float, 2>> vt;
        MeshBuilder builder;
        std::string raw;
        size_t lineno = 0;
        int objects = 0;
        while (std::getline(reader_, raw)) {
            ++lineno;
            if (!raw.empty() && raw.back() == '\r')
                raw.pop_back();
            for (char& c : raw)
                if (c == '\\')
                    c = ' ';
            size_t hash = raw.find('#');
            if (hash != std::string::npos)
                raw.resize(hash);
            std::istringstream ls(raw);
            std::vector<std::string> parts((std::istream_iterator<std::string>(ls)), std::istream_iterator<std::string>());
            if (parts.empty())
                continue;
            const std::string& tag = parts[0];
            if (tag == "o") {
                if (++objects > 1)
                    break;
            } else if (tag == "v") {
                v.push_back(Vector3(num(parts, 1, lineno), num(parts, 2, lineno), num(parts, 3, lineno)));
            } else if (tag == "vt") {
                vt.push_back({parts.size() > 1 ? num(parts, 1, lineno) : 0.0f,
                              parts.size() > 2 ? num(parts, 2, lineno) : 0.0f});
            } else if (tag == "vn") {
                vn.push_back(Vector3(num(parts, 1, lineno), num(parts, 2, lineno), num(parts, 3, lineno)));
            } else if (tag == "f") {
                std::vector<Corner> corners;
                for (size_t k = 1; k < parts.size(); ++k)
                    corners.push_back(corner(parts[k], v.size(), vt.size(), vn.size(), lineno));
                if (corners.size() < 3)
                    throw Error(BVHT_EINVAL, "line " + std::to_string(lineno) + ": a face needs three corners");
                for (size_t k = 1; k + 1 < corners.size(); ++k) {
                    const Corner* tri[3] = {&corners[0], &corners[k], &corners[k + 1]};
                    TextureCoordinates tc;
                    Normals nr;
                    for (int c = 0; c < 3; ++c) {
                        if (tri[c]->t >= 0)
                            tc.uv[c] = vt[tri[c]->t];
                        if (tri[c]->n >= 0)
                            nr.normals[c] = vn[tri[c]->n];
                    }
                    builder.with_primitive(Triangle(v[tri[0]->v], v[tri[1]->v], v[tri[2]->v]), tc, nr);
                }
            }
        }
        return builder.build();
    }

private:
    struct Corner {
        long v, t, n;
    };
    static float num(const std::vector<std::string>& parts, size_t i, size_t lineno)
    {
        if (i >= parts.size() || !detail::is_number(parts[i]))
            throw Error(BVHT_EINVAL, "line " + std::to_string(lineno) + ": expected a number");
        return detail::to_f32(parts[i]);
    }
    static long index(const std::string& s, size_t count, const std::string& tok, size_t lineno)
    {
        if (s.empty())
            return -1;
        char* end = nullptr;
        long k = std::strtol(s.c_str(), &end, 10);
        if (*end != '\0')
            throw Error(BVHT_EINVAL, "line " + std::to_string(lineno) + ": bad index in " + tok);
        k = k > 0 ? k - 1 : (long)count + k;
        if (k < 0 || k >= (long)count)
            throw Error(BVHT_EINVAL, "line " + std::to_string(lineno) + ": index " + s + " out of range in " + tok);
        return k;
    }
    static Corner corner(const std::string& tok, size_t nv, size_t nt, size_t nn, size_t lineno)
    {
        std::vector<std::string> f;
        size_t a = 0;
        while (true) {
            size_t b = tok.find('/', a);
            f.push_back(tok.substr(a, b == std::string::npos ? std::string::npos : b - a));
            if (b == std::string::npos)
                break;
            a = b + 1;
        }
        if (f.size() > 3 || f[0].empty())
            throw Error(BVHT_EINVAL, "line " + std::to_string(lineno) + ": bad face corner " + tok);
        return Corner{index(f[0], nv, tok, lineno), f.size() > 1 ? index(f[1], nt, tok, lineno) : -1,
                      f.size() > 2 ? index(f[2], nn, tok, lineno) : -1};
    }

    std::istream& reader_;
};

}  // namespace bvhtracer

#endif  // BVHTRACER_MESH_HPP
