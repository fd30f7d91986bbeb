// Wavefront OBJ mesh reader following e8util::wavefront_obj::load_geometry
// (resource.cpp:137-255).
//
// The reference classifies each line with std::regex_match against
//   v\s+(.+)\s+(.+)\s+(.+)\s*     vn\s+(.+)\s+(.+)\s+(.+)\s*
//   vt\s+(.+)\s+(.+)\s*           f\s+(.+)\s+(.+)\s+(.+)\s*
// (whole line, no leading blanks) and converts each group with std::stof / std::stoul. With
// greedy groups the last group is the last whitespace-separated token, the one before it the
// second-to-last token, and the first group everything from the first token up to there, of
// which std::stof reads the leading number. Lines that match none of the patterns (comments,
// o / s / usemtl lines, faces with fewer than three corners) are skipped. Each face group is
// split at '/' like std::getline (a trailing empty field disappears) and must give exactly three
// fields: position index (required), texture index, normal index (either may be empty), all
// 1-based and in range. Vertices keep their file order; each takes the normal and texture
// coordinate of the last face corner that names it.
//
// Deliberate differences: where the reference would index past its attribute arrays (a face
// corner without a normal or texture index) or abort on an exception it does not catch (a
// field that is not a number), this reader leaves the attribute unset or returns an error.

#include "objfile.h"

#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace x3d {

namespace {

struct Token {
    size_t begin, end; // [begin, end) in the line
};

// Whitespace-separated tokens after a keyword of length `kw`, if the line is the keyword
// followed by at least one whitespace character.
bool tokens_after(std::string const &line, char const *kw, std::vector<Token> &tok) {
    size_t k = 0;
    while (kw[k]) {
        if (k >= line.size() || line[k] != kw[k])
            return false;
        k++;
    }
    if (k >= line.size() || !std::isspace(static_cast<unsigned char>(line[k])))
        return false;
    tok.clear();
    size_t i = k;
    while (i < line.size()) {
        while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i])))
            i++;
        if (i >= line.size())
            break;
        size_t b = i;
        while (i < line.size() && !std::isspace(static_cast<unsigned char>(line[i])))
            i++;
        tok.push_back({b, i});
    }
    return true;
}

// std::stof / std::stoul of a field (leading blanks skipped, trailing characters ignored);
// false where they would throw.
bool to_float(std::string const &s, float *out) {
    try {
        *out = std::stof(s);
        return true;
    } catch (...) {
        return false;
    }
}

bool to_ulong(std::string const &s, unsigned long *out) {
    try {
        *out = std::stoul(s);
        return true;
    } catch (...) {
        return false;
    }
}

// The three groups of a pattern with three greedy groups: first token .. third-to-last token
// (as written), second-to-last token, last token.
void three_groups(std::string const &line, std::vector<Token> const &t, std::string g[3]) {
    size_t n = t.size();
    g[0] = line.substr(t[0].begin, t[n - 3].end - t[0].begin);
    g[1] = line.substr(t[n - 2].begin, t[n - 2].end - t[n - 2].begin);
    g[2] = line.substr(t[n - 1].begin, t[n - 1].end - t[n - 1].begin);
}

std::vector<std::string> split_fields(std::string const &s) {
    std::vector<std::string> parts;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, '/'))
        parts.push_back(item);
    return parts;
}

} // namespace

std::string load_obj(const char *path, ObjMesh &out) {
    std::ifstream file(path);
    if (!file.is_open())
        return std::string("cannot open ") + path;
    std::vector<float> pos, nrm, tex;                 // as read
    std::vector<uint32_t> iv, in, it;                 // per face corner
    std::vector<uint8_t> has_n, has_t;
    std::vector<Token> tok;
    std::string line;
    unsigned lineno = 0;
    auto err = [&](std::string const &what) { return "line " + std::to_string(lineno) + ": " + what; };
    while (std::getline(file, line)) {
        lineno++;
        std::string g[3];
        if (tokens_after(line, "v", tok) && tok.size() >= 3) {
            three_groups(line, tok, g);
            float x, y, z;
            if (!to_float(g[0], &x) || !to_float(g[1], &y) || !to_float(g[2], &z))
                return err("vertex is not a number");
            pos.insert(pos.end(), {x, y, z});
        } else if (tokens_after(line, "vn", tok) && tok.size() >= 3) {
            three_groups(line, tok, g);
            float x, y, z;
            if (!to_float(g[0], &x) || !to_float(g[1], &y) || !to_float(g[2], &z))
                return err("normal is not a number");
            nrm.insert(nrm.end(), {x, y, z});
        } else if (tokens_after(line, "vt", tok) && tok.size() >= 2) {
            size_t n = tok.size();
            std::string u = line.substr(tok[0].begin, tok[n - 2].end - tok[0].begin);
            std::string v = line.substr(tok[n - 1].begin, tok[n - 1].end - tok[n - 1].begin);
            float a, b;
            if (!to_float(u, &a) || !to_float(v, &b))
                return err("texture coordinate is not a number");
            tex.insert(tex.end(), {a, b});
        } else if (tokens_after(line, "f", tok) && tok.size() >= 3) {
            three_groups(line, tok, g);
            for (int c = 0; c < 3; c++) {
                std::vector<std::string> f = split_fields(g[c]);
                if (f.size() != 3)
                    return err("corner " + std::to_string(c + 1) + " does not have 3 attributes");
                unsigned long k;
                if (f[0].empty())
                    return err("vertex index is missing");
                if (!to_ulong(f[0], &k) || static_cast<uint32_t>(k - 1) >= pos.size() / 3)
                    return err("corner " + std::to_string(c + 1) + " names an illegal vertex");
                iv.push_back(static_cast<uint32_t>(k - 1));
                uint32_t ti = 0, ni = 0;
                bool ht = !f[1].empty(), hn = !f[2].empty();
                if (ht && (!to_ulong(f[1], &k) || static_cast<uint32_t>(k - 1) >= tex.size() / 2))
                    return err("corner " + std::to_string(c + 1) + " names an illegal texcoord");
                if (ht)
                    ti = static_cast<uint32_t>(k - 1);
                if (hn && (!to_ulong(f[2], &k) || static_cast<uint32_t>(k - 1) >= nrm.size() / 3))
                    return err("corner " + std::to_string(c + 1) + " names an illegal normal");
                if (hn)
                    ni = static_cast<uint32_t>(k - 1);
                it.push_back(ti);
                in.push_back(ni);
                has_t.push_back(ht ? 1 : 0);
                has_n.push_back(hn ? 1 : 0);
            }
        }
    }
    if (pos.empty())
        return "the mesh has no vertex data";
    const size_t nv = pos.size() / 3;
    out.verts = pos;
    out.normals.assign(3 * nv, 0.0f);
    out.uvs.assign(2 * nv, 0.0f);
    out.tris = iv;
    for (size_t c = 0; c < iv.size(); c++) {
        uint32_t v = iv[c];
        if (has_n[c])
            for (int a = 0; a < 3; a++)
                out.normals[3 * v + a] = nrm[3 * in[c] + a];
        if (has_t[c])
            for (int a = 0; a < 2; a++)
                out.uvs[2 * v + a] = tex[2 * it[c] + a];
    }
    return "";
}

} // namespace x3d
