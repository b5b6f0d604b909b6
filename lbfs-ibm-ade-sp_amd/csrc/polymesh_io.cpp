// OpenFOAM polyMesh ASCII reader and writer (the reference's mesh input,
// unstructured_mesh.cpp:670-941 and import_openfoam_BC_tags 2533-2608).
// Hexahedral meshes only: every face must be a quad.
#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace {

// strip the FoamFile header and comments, return the body
bool read_body(const std::string &path, std::string &body)
{
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str(), out;
    out.reserve(s.size());
    for (size_t i = 0; i < s.size(); i++) {
        if (s.compare(i, 2, "//") == 0) {
            while (i < s.size() && s[i] != '\n') i++;
            continue;
        }
        if (s.compare(i, 2, "/*") == 0) {
            size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? s.size() : e + 1;
            continue;
        }
        out.push_back(s[i]);
    }
    size_t h = out.find("FoamFile");
    if (h != std::string::npos) {
        size_t b = out.find('{', h), e = out.find('}', b);
        if (b != std::string::npos && e != std::string::npos) out.erase(h, e + 1 - h);
    }
    body = out;
    return true;
}

// tokens with parentheses as separate tokens
std::vector<std::string> tokens(const std::string &s)
{
    std::vector<std::string> t;
    std::string cur;
    for (char c : s) {
        if (c == '(' || c == ')' || c == '{' || c == '}' || c == ';') {
            if (!cur.empty()) { t.push_back(cur); cur.clear(); }
            t.push_back(std::string(1, c));
        } else if (std::isspace((unsigned char)c)) {
            if (!cur.empty()) { t.push_back(cur); cur.clear(); }
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) t.push_back(cur);
    return t;
}

bool read_points(const std::string &dir, std::vector<double> &p)
{
    std::string b;
    if (!read_body(dir + "/points", b)) return false;
    auto t = tokens(b);
    size_t i = 0;
    while (i < t.size() && !std::isdigit((unsigned char)t[i][0])) i++;
    if (i >= t.size()) return false;
    const long n = std::stol(t[i++]);
    p.clear();
    for (; i < t.size() && (long)p.size() < 3 * n; i++)
        if (t[i] != "(" && t[i] != ")") p.push_back(std::stod(t[i]));
    return (long)p.size() == 3 * n;
}

bool read_faces(const std::string &dir, std::vector<int> &fv, std::string &err)
{
    std::string b;
    if (!read_body(dir + "/faces", b)) { err = "cannot read faces"; return false; }
    // faces are "k(a b c ...)": split the leading count from its parenthesis
    std::string s;
    for (size_t i = 0; i < b.size(); i++) {
        s.push_back(b[i]);
        if (std::isdigit((unsigned char)b[i]) && i + 1 < b.size() && b[i + 1] == '(') s.push_back(' ');
    }
    auto t = tokens(s);
    size_t i = 0;
    while (i < t.size() && !std::isdigit((unsigned char)t[i][0])) i++;
    const long n = std::stol(t[i++]);
    if (t[i++] != "(") { err = "faces: malformed list"; return false; }
    fv.clear();
    for (long f = 0; f < n; f++) {
        const int k = std::stoi(t[i++]);
        if (k != 4) { err = "face " + std::to_string(f) + " is not a quad"; return false; }
        if (t[i++] != "(") { err = "faces: malformed face"; return false; }
        for (int j = 0; j < 4; j++) fv.push_back(std::stoi(t[i++]));
        i++;  // ')'
    }
    return true;
}

bool read_labels(const std::string &path, std::vector<int> &v)
{
    std::string b;
    if (!read_body(path, b)) return false;
    auto t = tokens(b);
    size_t i = 0;
    while (i < t.size() && !std::isdigit((unsigned char)t[i][0])) i++;
    if (i >= t.size()) return false;
    const long n = std::stol(t[i++]);
    v.clear();
    for (; i < t.size() && (long)v.size() < n; i++)
        if (t[i] != "(" && t[i] != ")") v.push_back(std::stoi(t[i]));
    return (long)v.size() == n;
}

struct Patch {
    std::string name, type;
    int n_faces = 0, start = 0;
};

bool read_boundary(const std::string &dir, std::vector<Patch> &patches)
{
    std::string b;
    if (!read_body(dir + "/boundary", b)) return false;
    auto t = tokens(b);
    size_t i = 0;
    while (i < t.size() && !std::isdigit((unsigned char)t[i][0])) i++;
    if (i >= t.size()) return false;
    const int n = std::stoi(t[i++]);
    patches.clear();
    while (i < t.size() && (int)patches.size() < n) {
        if (t[i] == "(" || t[i] == ")") { i++; continue; }
        Patch p;
        p.name = t[i++];
        if (i >= t.size() || t[i] != "{") return false;
        i++;
        while (i < t.size() && t[i] != "}") {
            const std::string key = t[i++];
            std::string val;
            while (i < t.size() && t[i] != ";" && t[i] != "}") val += (val.empty() ? "" : " ") + t[i++];
            if (i < t.size() && t[i] == ";") i++;
            if (key == "type") p.type = val;
            else if (key == "nFaces") p.n_faces = std::stoi(val);
            else if (key == "startFace") p.start = std::stoi(val);
        }
        i++;
        patches.push_back(p);
    }
    return (int)patches.size() == n;
}

}  // namespace

extern "C" int lbfs_polymesh_counts(const char *dir, int *n_points, int *n_cells, int *n_faces, int *n_internal,
                                    int *n_patches)
{
    if (!dir || !n_points || !n_cells || !n_faces || !n_internal || !n_patches) return LBFS_E_INVALID_ARG;
    std::vector<double> p;
    std::vector<int> own, nb;
    std::vector<Patch> pt;
    if (!read_points(dir, p) || !read_labels(std::string(dir) + "/owner", own) ||
        !read_labels(std::string(dir) + "/neighbour", nb) || !read_boundary(dir, pt))
        return LBFS_E_INVALID_ARG;
    int nc = 0;
    for (int o : own) nc = o + 1 > nc ? o + 1 : nc;
    for (int o : nb) nc = o + 1 > nc ? o + 1 : nc;
    *n_points = (int)p.size() / 3;
    *n_cells = nc;
    *n_faces = (int)own.size();
    *n_internal = (int)nb.size();
    *n_patches = (int)pt.size();
    return LBFS_OK;
}

extern "C" int lbfs_polymesh_read(const char *dir, double *points, int *face_vertices, int *owner, int *neighbour,
                                  int *face_patch, double *face_shift)
{
    if (!dir || !points || !face_vertices || !owner || !neighbour || !face_patch) return LBFS_E_INVALID_ARG;
    std::vector<double> p;
    std::vector<int> fv, own, nb;
    std::vector<Patch> pt;
    std::string err;
    if (!read_points(dir, p) || !read_faces(dir, fv, err) || !read_labels(std::string(dir) + "/owner", own) ||
        !read_labels(std::string(dir) + "/neighbour", nb) || !read_boundary(dir, pt))
        return LBFS_E_INVALID_ARG;
    std::copy(p.begin(), p.end(), points);
    std::copy(fv.begin(), fv.end(), face_vertices);
    std::copy(own.begin(), own.end(), owner);
    std::copy(nb.begin(), nb.end(), neighbour);
    for (size_t f = 0; f < own.size(); f++) face_patch[f] = -1;
    for (size_t k = 0; k < pt.size(); k++)
        for (int f = pt[k].start; f < pt[k].start + pt[k].n_faces && f < (int)own.size(); f++) face_patch[f] = (int)k;
    if (face_shift) {
        // periodic wrap faces are internal faces with a translation of the
        // neighbour image, kept in an extra vectorField next to the polyMesh
        std::vector<double> sh;
        for (size_t f = 0; f < 3 * own.size(); f++) face_shift[f] = 0.0;
        std::string b;
        if (read_body(std::string(dir) + "/lbfsFaceShift", b)) {
            auto t = tokens(b);
            size_t i = 0;
            while (i < t.size() && !std::isdigit((unsigned char)t[i][0])) i++;
            const long n = i < t.size() ? std::stol(t[i++]) : 0;
            std::vector<double> v;
            for (; i < t.size() && (long)v.size() < 3 * n; i++)
                if (t[i] != "(" && t[i] != ")") v.push_back(std::stod(t[i]));
            for (size_t f = 0; f < v.size() && f < 3 * own.size(); f++) face_shift[f] = v[f];
        }
    }
    return LBFS_OK;
}

extern "C" int lbfs_polymesh_patch(const char *dir, int index, char *name, int name_len, char *type, int type_len)
{
    if (!dir || !name || !type || name_len <= 0 || type_len <= 0) return LBFS_E_INVALID_ARG;
    std::vector<Patch> pt;
    if (!read_boundary(dir, pt) || index < 0 || index >= (int)pt.size()) return LBFS_E_INVALID_ARG;
    std::snprintf(name, name_len, "%s", pt[index].name.c_str());
    std::snprintf(type, type_len, "%s", pt[index].type.c_str());
    return LBFS_OK;
}

extern "C" int lbfs_polymesh_write(const char *dir, int n_points, const double *points, int n_cells, int n_faces,
                                   int n_internal, const int *face_vertices, const int *owner, const int *neighbour,
                                   int n_patches, const char *const *patch_names, const char *const *patch_types,
                                   const int *patch_start, const int *patch_size, const double *face_shift)
{
    if (!dir || !points || !face_vertices || !owner || (n_internal > 0 && !neighbour) || n_patches < 0)
        return LBFS_E_INVALID_ARG;
    const std::string d(dir);
    auto header = [](std::ostream &o, const char *cls, const char *obj, const std::string &note) {
        o << "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       " << cls
          << ";\n";
        if (!note.empty()) o << "    note        \"" << note << "\";\n";
        o << "    location    \"constant/polyMesh\";\n    object      " << obj << ";\n}\n\n";
    };
    const std::string note = "nPoints:" + std::to_string(n_points) + "  nCells:" + std::to_string(n_cells) +
                             "  nFaces:" + std::to_string(n_faces) + "  nInternalFaces:" + std::to_string(n_internal);
    {
        std::ofstream o(d + "/points");
        if (!o) return LBFS_E_INVALID_ARG;
        header(o, "vectorField", "points", "");
        o.precision(17);
        o << n_points << "\n(\n";
        for (int i = 0; i < n_points; i++) o << "(" << points[3 * i] << " " << points[3 * i + 1] << " " << points[3 * i + 2] << ")\n";
        o << ")\n";
    }
    {
        std::ofstream o(d + "/faces");
        header(o, "faceList", "faces", "");
        o << n_faces << "\n(\n";
        for (int f = 0; f < n_faces; f++)
            o << "4(" << face_vertices[4 * f] << " " << face_vertices[4 * f + 1] << " " << face_vertices[4 * f + 2] << " "
              << face_vertices[4 * f + 3] << ")\n";
        o << ")\n";
    }
    {
        std::ofstream o(d + "/owner");
        header(o, "labelList", "owner", note);
        o << n_faces << "\n(\n";
        for (int f = 0; f < n_faces; f++) o << owner[f] << "\n";
        o << ")\n";
    }
    {
        std::ofstream o(d + "/neighbour");
        header(o, "labelList", "neighbour", note);
        o << n_internal << "\n(\n";
        for (int f = 0; f < n_internal; f++) o << neighbour[f] << "\n";
        o << ")\n";
    }
    {
        std::ofstream o(d + "/boundary");
        header(o, "polyBoundaryMesh", "boundary", "");
        o << n_patches << "\n(\n";
        for (int k = 0; k < n_patches; k++)
            o << "    " << patch_names[k] << "\n    {\n        type            " << patch_types[k]
              << ";\n        nFaces          " << patch_size[k] << ";\n        startFace       " << patch_start[k]
              << ";\n    }\n";
        o << ")\n";
    }
    if (face_shift) {
        std::ofstream o(d + "/lbfsFaceShift");
        header(o, "vectorField", "lbfsFaceShift", "");
        o.precision(17);
        o << n_faces << "\n(\n";
        for (int f = 0; f < n_faces; f++)
            o << "(" << face_shift[3 * f] << " " << face_shift[3 * f + 1] << " " << face_shift[3 * f + 2] << ")\n";
        o << ")\n";
    }
    return LBFS_OK;
}
