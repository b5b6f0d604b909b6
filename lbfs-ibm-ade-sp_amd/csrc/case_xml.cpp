// Case files in the reference's XML schema (preprocessor.cpp:87-857): the
// parts the time step reads — global_variables, geometry, initial_conditions,
// uns_boundary_conditions and lagrangian_objects.  A small element-tree
// parser of its own (elements and text; attributes, comments, declarations
// and CDATA-free documents), no external library.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace {

struct Node {
    std::string name, text;
    std::vector<std::unique_ptr<Node>> kids;
    const Node *child(const char *n) const
    {
        for (auto &k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
};

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// recursive-descent parse of one element starting at '<' (pos points at it)
bool parse_element(const std::string &s, size_t &pos, Node &out, std::string &err)
{
    size_t e = s.find('>', pos);
    if (e == std::string::npos) { err = "unterminated tag"; return false; }
    std::string tag = s.substr(pos + 1, e - pos - 1);
    const bool empty = !tag.empty() && tag.back() == '/';
    if (empty) tag.pop_back();
    size_t sp = 0;
    while (sp < tag.size() && !std::isspace((unsigned char)tag[sp])) sp++;
    out.name = tag.substr(0, sp);
    pos = e + 1;
    if (empty) return true;
    std::string text;
    while (pos < s.size()) {
        if (s.compare(pos, 4, "<!--") == 0) {
            size_t c = s.find("-->", pos);
            if (c == std::string::npos) { err = "unterminated comment"; return false; }
            pos = c + 3;
        } else if (s.compare(pos, 2, "</") == 0) {
            size_t c = s.find('>', pos);
            if (c == std::string::npos) { err = "unterminated closing tag"; return false; }
            if (trim(s.substr(pos + 2, c - pos - 2)) != out.name) { err = "mismatched closing tag of " + out.name; return false; }
            pos = c + 1;
            out.text = trim(text);
            return true;
        } else if (s[pos] == '<') {
            auto k = std::make_unique<Node>();
            if (!parse_element(s, pos, *k, err)) return false;
            out.kids.push_back(std::move(k));
        } else {
            text.push_back(s[pos++]);
        }
    }
    err = "unterminated element " + out.name;
    return false;
}

bool parse_document(const char *path, Node &root, std::string &err)
{
    std::ifstream f(path);
    if (!f) { err = "cannot open file"; return false; }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    size_t pos = 0;
    while (pos < s.size()) {
        if (s.compare(pos, 2, "<?") == 0) {
            pos = s.find("?>", pos);
            if (pos == std::string::npos) break;
            pos += 2;
        } else if (s.compare(pos, 4, "<!--") == 0) {
            pos = s.find("-->", pos);
            if (pos == std::string::npos) break;
            pos += 3;
        } else if (s[pos] == '<') {
            return parse_element(s, pos, root, err);
        } else {
            pos++;
        }
    }
    err = "no root element";
    return false;
}

double num(const Node *p, const char *name, double dflt, bool *found = nullptr)
{
    const Node *c = p ? p->child(name) : nullptr;
    if (found) *found = c != nullptr;
    if (!c || c->text.empty()) return dflt;
    return std::strtod(c->text.c_str(), nullptr);
}

std::string str(const Node *p, const char *name)
{
    const Node *c = p ? p->child(name) : nullptr;
    return c ? c->text : std::string();
}

// vel/rho type names (preprocessor.cpp:469-491)
int bc_type(const std::string &t)
{
    static const char *names[] = {"dirichlet", "neumann", "periodic", "parabolic-N", "parabolic-W",
                                  "inlet", "wall", "symmetry-x", "parabolic", "shear"};
    for (int k = 0; k < 10; k++)
        if (t == names[k]) return k + 1;
    if (t == "empty") return LBFS_BC_EMPTY;
    return 0;
}

void copy_str(char *dst, size_t n, const std::string &s) { std::snprintf(dst, n, "%s", s.c_str()); }

}  // namespace

extern "C" int lbfs_case_read(const char *path, lbfs_case *c)
{
    if (!path || !c) return LBFS_E_INVALID_ARG;
    Node root;
    std::string err;
    std::memset(c, 0, sizeof(*c));
    if (!parse_document(path, root, err)) {
        copy_str(c->error, sizeof(c->error), err);
        return LBFS_E_INVALID_ARG;
    }
    const Node *gv = root.child("global_variables");
    const Node *geo = root.child("geometry");
    const Node *ic = root.child("initial_conditions");
    if (!gv || !geo) {
        copy_str(c->error, sizeof(c->error), "global_variables and geometry are required");
        return LBFS_E_INVALID_ARG;
    }
    // global_variables (preprocessor.cpp:795-857) and their derived values
    // (global_variables.cpp:20-68): nu = U L / Re, U read from mach_no
    const double U = num(gv, "mach_no", 0.0), Re = num(gv, "reynolds_no", 1.0);
    const double scale = num(gv, "scale", 1.0);
    c->max_velocity = U;
    c->reynolds_number = Re;
    c->time_marching_step = num(gv, "time_marching_step", 0.0);
    c->simulation_length = num(gv, "simulation_length", 0.0) * scale;
    c->output_step = (int)num(gv, "output_step", 0.0);
    c->fneq_min = num(gv, "fneq_min", 0.0);
    c->scale = scale;
    double gamma = num(gv, "pre_condition_gamma", 1.0);
    const bool precond = str(gv, "preconditioning") == "true";
    if (precond && std::fabs(gamma - 1.0) < 0.00001) gamma = U * 2.0 * std::sqrt(3.0);
    c->gamma = gamma;
    const std::string ts = str(gv, "time_stepping");
    c->time_stepping = ts == "local" ? LBFS_TS_LOCAL : ts == "min" ? LBFS_TS_MIN : ts == "constant" ? LBFS_TS_CONSTANT : 0;
    const std::string gc = str(gv, "gradient_calc");
    c->gradient_type = gc == "LS" ? LBFS_GRAD_LS : gc == "GG" ? LBFS_GRAD_GG : LBFS_GRAD_HYBRID;
    c->mesh_type = (int)num(gv, "mesh_type", 0.0);
    c->gpu_solver = (int)num(gv, "gpu_solver", 0.0);
    c->gpu_device = (int)num(gv, "gpu_device", 0.0);
    copy_str(c->import_file, sizeof(c->import_file), str(gv, "import_file"));
    copy_str(c->import_format, sizeof(c->import_format), str(gv, "import_format"));
    // geometry (preprocessor.cpp:743-776)
    c->domain[0] = num(geo, "x", 0.0);
    c->domain[1] = num(geo, "y", 0.0);
    c->domain[2] = num(geo, "z", 0.0);
    c->spacing[0] = num(geo, "dx", 0.0);
    c->spacing[1] = num(geo, "dy", 0.0);
    c->spacing[2] = num(geo, "dz", 0.0);
    c->streaming_dt = num(geo, "streaming_dt", 0.0);
    c->origin[0] = num(geo, "origin_x", 0.0);
    c->origin[1] = num(geo, "origin_y", 0.0);
    c->origin[2] = num(geo, "origin_z", 0.0);
    c->ref_length = num(geo, "ref_length", 0.0);
    c->visc = U * c->ref_length / Re;
    if ((int)num(gv, "testcase", 0.0) == 8) c->visc = 0.0;
    // initial conditions (preprocessor.cpp:670-738): signed body force
    c->average_rho = num(ic, "average_rho", 1.0);
    c->initial_velocity[0] = num(ic, "u", 0.0);
    c->initial_velocity[1] = num(ic, "v", 0.0);
    c->initial_velocity[2] = num(ic, "w", 0.0);
    const double pg = num(ic, "pressure_gradient", 0.0);
    const std::string pd = str(ic, "pressure_direction");
    // the GPU solver applies +pressure_gradient along pressure_direction and
    // nothing without one (Solver_gpu.cu:834-839)
    if (pd == "x" || pd == "y" || pd == "z") c->body_force[pd == "x" ? 0 : pd == "y" ? 1 : 2] = pg;
    // boundary conditions by patch name (preprocessor.cpp:419-530); velocity
    // values scale by max_velocity / scale as there
    const Node *ub = root.child("uns_boundary_conditions");
    const double factor = U / scale;
    if (ub)
        for (auto &k : ub->kids) {
            if (k->name != "bc") continue;
            if (c->n_patches >= LBFS_CASE_MAX_PATCHES) {
                copy_str(c->error, sizeof(c->error), "too many boundary conditions");
                return LBFS_E_UNSUPPORTED;
            }
            lbfs_case_patch &p = c->patches[c->n_patches++];
            copy_str(p.name, sizeof(p.name), str(k.get(), "name"));
            p.rho_type = bc_type(str(k.get(), "rho_type"));
            p.vel_type = bc_type(str(k.get(), "vel_type"));
            p.values[0] = num(k.get(), "rho", 0.0);
            p.values[1] = num(k.get(), "u", 0.0) * factor;
            p.values[2] = num(k.get(), "v", 0.0) * factor;
            p.values[3] = num(k.get(), "w", 0.0) * factor;
        }
    // lagrangian objects (preprocessor.cpp:87-414).  As there, one object
    // record carries over from one <object> to the next, so fields an object
    // leaves out keep the previous object's value, and <sphere_reference>
    // sets the periodic flag while sphere_reference itself stays true
    // (lagrangian_object.h:41).
    const Node *lo = root.child("lagrangian_objects");
    lbfs_case_object cur;
    std::memset(&cur, 0, sizeof(cur));
    cur.internal_viscosity_ratio = 5.0;
    cur.volume_ratio = 1.0;
    cur.sphere_reference = 1;
    if (lo)
        for (auto &k : lo->kids) {
            if (k->name != "object") continue;
            if (c->n_objects >= LBFS_CASE_MAX_OBJECTS) {
                copy_str(c->error, sizeof(c->error), "too many lagrangian objects");
                return LBFS_E_UNSUPPORTED;
            }
            const Node *o = k.get();
            copy_str(cur.name, sizeof(cur.name), str(o, "name"));
            cur.num_nodes = (int)num(o, "num_nodes", cur.num_nodes);
            cur.radius = num(o, "radius", cur.radius);
            cur.centre[0] = num(o, "centre_x", cur.centre[0]);
            cur.centre[1] = num(o, "centre_y", cur.centre[1]);
            cur.centre[2] = num(o, "centre_z", cur.centre[2]);
            const std::string ty = str(o, "type");
            cur.type = ty == "rigid_cylinder" ? 1 : ty == "spring_network" ? 2 : 0;
            bool f;
            std::string s;
            if (!(s = str(o, "mesh_file")).empty()) copy_str(cur.mesh_file, sizeof(cur.mesh_file), s);
            if (!(s = str(o, "stress_free_mesh")).empty()) copy_str(cur.stress_free_mesh, sizeof(cur.stress_free_mesh), s);
            if (!(s = str(o, "curvature_mesh")).empty()) copy_str(cur.curvature_mesh, sizeof(cur.curvature_mesh), s);
            double v;
#define LBFS_OPT(field, tag) v = num(o, tag, 0.0, &f); if (f) cur.field = v;
            LBFS_OPT(local_area_modulus, "local_area_modulus")
            LBFS_OPT(global_area_modulus, "global_area_modulus")
            LBFS_OPT(volume_modulus, "volume_modulus")
            LBFS_OPT(shear_modulus, "shear_modulus")
            LBFS_OPT(local_bending_modulus, "local_bending_modulus")
            LBFS_OPT(global_bending_modulus, "global_bending_modulus")
            LBFS_OPT(viscosity, "viscosity")
            LBFS_OPT(wlc_ratio, "wlc_ratio")
            LBFS_OPT(internal_viscosity_ratio, "internal_viscosity_ratio")
            LBFS_OPT(volume_ratio, "volume_ratio")
            LBFS_OPT(curvature_ratio, "curvature_ratio")
#undef LBFS_OPT
            cur.periodic = str(o, "periodic") == "true" ? 1 : 0;
            cur.periodic = str(o, "sphere_reference") == "true" ? 1 : 0;
            cur.mz_importer = str(o, "mz_importer") == "true" ? 1 : 0;
            c->objects[c->n_objects++] = cur;
        }
    return LBFS_OK;
}
