// Handle-based C API (include/cedar_capi.h): configs, topologies, vectors, matrices and solvers
// addressed by int handles, over the device-pointer API of include/cedar_hip.h.
//
// Objects live in one process-wide slot table; a handle is (kind << 28) | (slot + 1), as in the
// reference (include/cedar/interface/object.h:8-30, src/interface/object.cc).  Vectors and
// matrices keep host arrays in the solver's layout (Fortran order, one ghost layer); matvec and
// solves copy them to HBM, run the HIP kernels, and copy results back.
#include "../../include/cedar_capi.h"
#include "../../include/cedar_hip.h"
#include "cedar_hip.hpp"

#include <hip/hip_runtime.h>

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------------------------------------
// JSON values and a recursive-descent parser (RFC 8259; no comments, no trailing commas)
// ------------------------------------------------------------------------------------------
struct jvalue {
    enum kind_t { NUL, BOOL, INT, REAL, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    long long i = 0;
    double r = 0.0;
    std::string s;
    std::vector<jvalue> arr;
    std::vector<std::pair<std::string, jvalue>> obj;

    const jvalue *member(const std::string &key) const {
        if (kind != OBJ) return nullptr;
        for (auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    // dotted path ("solver.cycle.nrelax-pre"); "" is the value itself
    const jvalue *at(const std::string &path) const {
        const jvalue *v = this;
        std::size_t pos = 0;
        while (v && pos <= path.size() && !path.empty()) {
            std::size_t dot = path.find('.', pos);
            std::string key = path.substr(pos, dot == std::string::npos ? std::string::npos
                                                                        : dot - pos);
            v = v->member(key);
            if (dot == std::string::npos) break;
            pos = dot + 1;
        }
        return v;
    }
};

class parser {
  public:
    explicit parser(const std::string &t) : t_(t) {}
    bool parse(jvalue &out) {
        ws();
        if (!value(out, 0)) return false;
        ws();
        return p_ == t_.size();
    }

  private:
    const std::string &t_;
    std::size_t p_ = 0;

    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' ||
                                  t_[p_] == '\r'))
            ++p_;
    }
    bool lit(const char *w) {
        std::size_t n = std::strlen(w);
        if (t_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }
    static void utf8(std::string &o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(unsigned &v) {
        if (p_ + 4 > t_.size()) return false;
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = t_[p_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string &o) {
        if (p_ >= t_.size() || t_[p_] != '"') return false;
        ++p_;
        while (p_ < t_.size()) {
            char c = t_[p_++];
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return false;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p_ >= t_.size()) return false;
            char e = t_[p_++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                unsigned cp;
                if (!hex4(cp)) return false;
                if (cp >= 0xD800 && cp < 0xDC00) {
                    unsigned lo;
                    if (!lit("\\u") || !hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return false;
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(o, cp);
                break;
            }
            default: return false;
            }
        }
        return false;
    }
    bool number(jvalue &v) {
        std::size_t s = p_;
        bool real = false;
        if (p_ < t_.size() && t_[p_] == '-') ++p_;
        if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) return false;
        if (t_[p_] == '0') ++p_;
        else
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        if (p_ < t_.size() && t_[p_] == '.') {
            real = true;
            ++p_;
            if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) return false;
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        }
        if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
            real = true;
            ++p_;
            if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
            if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) return false;
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        }
        std::string txt = t_.substr(s, p_ - s);
        if (!real) {
            errno = 0;
            long long x = std::strtoll(txt.c_str(), nullptr, 10);
            if (errno == 0) {
                v.kind = jvalue::INT;
                v.i = x;
                v.r = (double)x;
                return true;
            }
        }
        v.kind = jvalue::REAL;
        v.r = std::strtod(txt.c_str(), nullptr);
        return true;
    }
    bool value(jvalue &v, int depth) {
        if (depth > 256 || p_ >= t_.size()) return false;
        char c = t_[p_];
        if (c == '{') {
            ++p_;
            v.kind = jvalue::OBJ;
            ws();
            if (p_ < t_.size() && t_[p_] == '}') {
                ++p_;
                return true;
            }
            for (;;) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p_ >= t_.size() || t_[p_] != ':') return false;
                ++p_;
                ws();
                jvalue m;
                if (!value(m, depth + 1)) return false;
                // a repeated key keeps the last value, as nlohmann::json does
                bool found = false;
                for (auto &kv : v.obj)
                    if (kv.first == key) {
                        kv.second = std::move(m);
                        found = true;
                        break;
                    }
                if (!found) v.obj.emplace_back(std::move(key), std::move(m));
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == '}') {
                    ++p_;
                    return true;
                }
                return false;
            }
        }
        if (c == '[') {
            ++p_;
            v.kind = jvalue::ARR;
            ws();
            if (p_ < t_.size() && t_[p_] == ']') {
                ++p_;
                return true;
            }
            for (;;) {
                ws();
                jvalue m;
                if (!value(m, depth + 1)) return false;
                v.arr.push_back(std::move(m));
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == ']') {
                    ++p_;
                    return true;
                }
                return false;
            }
        }
        if (c == '"') {
            v.kind = jvalue::STR;
            return string(v.s);
        }
        if (lit("true")) {
            v.kind = jvalue::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = jvalue::BOOL;
            return true;
        }
        if (lit("null")) {
            v.kind = jvalue::NUL;
            return true;
        }
        return number(v);
    }
};

// typed reads with the conversions of nlohmann::json's get<T>() (include/cedar/config.h:80-101):
// numbers convert among themselves, booleans convert to numbers, nothing converts to a string
bool as_int(const jvalue *v, int &out) {
    if (!v) return false;
    if (v->kind == jvalue::INT) out = (int)v->i;
    else if (v->kind == jvalue::REAL) out = (int)v->r;
    else if (v->kind == jvalue::BOOL) out = v->b ? 1 : 0;
    else return false;
    return true;
}
bool as_real(const jvalue *v, double &out) {
    if (!v) return false;
    if (v->kind == jvalue::INT) out = (double)v->i;
    else if (v->kind == jvalue::REAL) out = v->r;
    else if (v->kind == jvalue::BOOL) out = v->b ? 1.0 : 0.0;
    else return false;
    return true;
}
bool as_bool(const jvalue *v, bool &out) {
    if (!v || v->kind != jvalue::BOOL) return false;
    out = v->b;
    return true;
}
bool as_string(const jvalue *v, std::string &out) {
    if (!v || v->kind != jvalue::STR) return false;
    out = v->s;
    return true;
}
template <typename T, typename F>
T get_or(const jvalue &root, const char *path, T dflt, F conv) {
    T out;
    return conv(root.at(path), out) ? out : dflt;
}

// ------------------------------------------------------------------------------------------
// logging: levels of src/util/log.cc:53-95, lines as LogLevelBuf::sync writes them (:130-146)
// ------------------------------------------------------------------------------------------
enum : unsigned {
    LOG_STATUS = 2u << 0, LOG_INFO = 2u << 1, LOG_ERROR = 2u << 2, LOG_MEMORY = 2u << 3,
    LOG_DEBUG = 2u << 4, LOG_TIMER = 2u << 5
};
unsigned g_log_level = LOG_STATUS | LOG_ERROR;

void log_line(unsigned lvl, const std::string &msg) {
    if (!(g_log_level & lvl)) return;
    std::time_t t = std::time(nullptr);
    std::tm tm = *std::localtime(&t);
    char date[64];
    std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y", &tm);
    const char *color = lvl == LOG_ERROR ? "\033[31m" : lvl == LOG_MEMORY ? "\033[34m"
                      : lvl == LOG_DEBUG ? "\033[35m" : "\033[39m";
    std::ostream &os = lvl == LOG_ERROR ? std::cerr : std::cout;
    os << "\033[32m[Cedar <" << date << ">] \033[39m" << color << msg << "\n\033[39m";
    os.flush();
}

// ------------------------------------------------------------------------------------------
// objects and the handle table
// ------------------------------------------------------------------------------------------
enum kind_t { KTOPO = 1, KSOLVER = 2, KMAT = 3, KVEC = 4, KCONFIG = 5 };
constexpr int KIND_SHIFT = 28, KEY_MASK = 0x0fffffff;

struct config_obj {
    jvalue root;
};
struct topo_obj {
    int nd;
    unsigned n[3];   // interior points per dimension
};
struct vec_obj {
    int nd;
    int topo;
    std::shared_ptr<topo_obj> grid;
    std::vector<double> data;   // ghost-inclusive, Fortran order
    std::size_t len[3];
};
struct dev_buf {
    double *p = nullptr;
    std::size_t n = 0;
    ~dev_buf() {
        if (p) (void)hipFree(p);
    }
    bool reserve(std::size_t count) {
        if (n >= count) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, count * sizeof(double)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
};
struct mat_obj {
    int nd;
    int topo, conf;
    std::shared_ptr<topo_obj> grid;
    std::shared_ptr<config_obj> cfg;
    int nstncl;
    std::size_t len[3];
    std::size_t npts;
    std::vector<double> so;   // nstncl planes of npts
    dev_buf dso, dx, dy;
    bool dirty = true;
};
struct solver_obj {
    int nd;
    int maxiter;
    int pmask = 0;   // periodic directions (per_mask)
    cdr_hip_solver2d *s2 = nullptr;
    cdr_hip_solver3d *s3 = nullptr;
    std::size_t npts;
    std::size_t len[3];
    dev_buf dso, dx, db;
    ~solver_obj() {
        if (s2) cdr_hip_solver2d_destroy(s2);
        if (s3) cdr_hip_solver3d_destroy(s3);
    }
};

struct slot {
    int kind = 0;
    int refcount = 0;
    bool released = false;   // its handle was freed while still referenced
    std::shared_ptr<void> obj;
};
// The table is never destroyed: objects the caller did not free would otherwise release device
// memory during static destruction, after the HIP runtime may already be gone.
std::vector<slot> &g_slots = *new std::vector<slot>();

template <typename T>
int create(int kind, std::shared_ptr<T> obj) {
    slot s;
    s.kind = kind;
    s.obj = std::static_pointer_cast<void>(obj);
    g_slots.push_back(std::move(s));
    return (kind << KIND_SHIFT) | (int)g_slots.size();
}
slot *lookup(int handle, int kind) {
    if (((unsigned)handle >> KIND_SHIFT) != (unsigned)kind) return nullptr;
    std::size_t k = (std::size_t)(handle & KEY_MASK);
    if (k == 0 || k > g_slots.size()) return nullptr;
    slot &s = g_slots[k - 1];
    if (s.kind != kind || !s.obj || s.released) return nullptr;
    return &s;
}
template <typename T>
std::shared_ptr<T> get(int handle, int kind) {
    slot *s = lookup(handle, kind);
    return s ? std::static_pointer_cast<T>(s->obj) : nullptr;
}
void incref(int handle) {
    std::size_t k = (std::size_t)(handle & KEY_MASK);
    if (k >= 1 && k <= g_slots.size()) g_slots[k - 1].refcount++;
}
void decref(int handle) {
    std::size_t k = (std::size_t)(handle & KEY_MASK);
    if (k < 1 || k > g_slots.size()) return;
    slot &s = g_slots[k - 1];
    if (s.refcount > 0) s.refcount--;
    if (s.refcount == 0 && s.released) s.obj.reset();
}
// frees the handle; an object still referenced (a topology or config used by a matrix or
// vector) lives on until its last user is freed
int release(int *handle, int kind, int err, int null_handle) {
    slot *s = lookup(*handle, kind);
    if (!s) return err;
    if (s->refcount == 0) s->obj.reset();
    else s->released = true;
    *handle = null_handle;
    return CEDAR_SUCCESS;
}

bool valid_topo_dims(const unsigned *n, int nd) {
    long long pts = 1;
    for (int d = 0; d < nd; ++d) {
        if (n[d] < 1 || n[d] > (1u << 30)) return false;
        pts *= (long long)n[d] + 2;
    }
    return pts < (1LL << 31);
}

// ------------------------------------------------------------------------------------------
// settings from a config (src/multilevel_settings.cc:15-61, src/kernel_params.cc:72-89)
// ------------------------------------------------------------------------------------------
const std::map<std::string, int> kRelax = {
    {"point", CDR_RELAX_POINT},       {"line-x", CDR_RELAX_LINE_X},
    {"line-y", CDR_RELAX_LINE_Y},     {"line-xy", CDR_RELAX_LINE_XY},
    {"plane-xy", CDR_RELAX_PLANE_XY}, {"plane-xz", CDR_RELAX_PLANE_XZ},
    {"plane-yz", CDR_RELAX_PLANE_YZ}, {"plane-xyz", CDR_RELAX_PLANE_XYZ}};

struct ml_keys {
    int relaxation, cycle, nrelax_pre, nrelax_post, num_levels, maxiter, min_coarse;
    double tol;
    std::string relax_name, cg;
};

bool read_ml(const jvalue &root, ml_keys &k, std::string &err) {
    k.relax_name = get_or<std::string>(root, "solver.relaxation", "point", as_string);
    auto r = kRelax.find(k.relax_name);
    if (r == kRelax.end()) {
        err = "invalid relaxation type: " + k.relax_name;
        return false;
    }
    k.relaxation = r->second;
    std::string cyc = get_or<std::string>(root, "solver.cycle.type", "v", as_string);
    if (cyc != "v" && cyc != "f") {
        err = "invalid cycle type: " + cyc;
        return false;
    }
    k.cycle = cyc == "v" ? CDR_CYCLE_V : CDR_CYCLE_F;
    k.nrelax_pre = get_or(root, "solver.cycle.nrelax-pre", 2, as_int);
    k.nrelax_post = get_or(root, "solver.cycle.nrelax-post", 1, as_int);
    k.num_levels = get_or(root, "solver.num-levels", -1, as_int);
    k.maxiter = get_or(root, "solver.max-iter", 10, as_int);
    k.tol = get_or(root, "solver.tol", 1e-8, as_real);
    k.min_coarse = get_or(root, "solver.min_coarse", 3, as_int);
    k.cg = get_or<std::string>(root, "solver.cg-solver", "LU", as_string);
    if (k.cg != "LU" && k.cg != "cedar" && k.cg != "redist") {
        err = "invalid value for solver.cg-solver";
        return false;
    }
    return true;
}

void to_settings(const ml_keys &k, const jvalue &root, cdr_solver_settings &s) {
    cdr_hip_default_settings(&s);
    s.relaxation = k.relaxation;
    s.nrelax_pre = k.nrelax_pre;
    s.nrelax_post = k.nrelax_post;
    s.num_levels = k.num_levels;
    s.maxiter = k.maxiter;
    s.tol = k.tol;
    s.min_coarse = k.min_coarse;
    s.cycle = k.cycle;
    s.graph = get_or(root, "gpu.graph", true, as_bool) ? 1 : 0;
    s.fused_coarse = get_or(root, "gpu.fused-coarse", true, as_bool) ? 1 : 0;
    s.stall_cycles = std::max(0, get_or(root, "gpu.stall-cycles", 0, as_int));
}

bool read_precision(const jvalue &root, int &prec, std::string &err) {
    std::string p = get_or<std::string>(root, "gpu.precision", "f64", as_string);
    if (p == "f64") prec = CDR_PREC_F64;
    else if (p == "mixed") prec = CDR_PREC_MIXED;
    else {
        err = "invalid gpu.precision: " + p;
        return false;
    }
    return true;
}

// grid.periodic as Cedar's per_mask (src/kernel_params.cc:74-76, include/cedar/kernel_params.h
// :35-42): bit d set when entry d is true
int periodic_mask(const jvalue &root) {
    const jvalue *p = root.at("grid.periodic");
    if (!p || p->kind != jvalue::ARR) return 0;
    int mask = 0;
    for (std::size_t d = 0; d < p->arr.size() && d < 3; ++d)
        if (p->arr[d].kind == jvalue::BOOL && p->arr[d].b) mask |= 1 << d;
    return mask;
}

// ml_settings operator<< (src/multilevel_settings.cc:63-95)
std::string settings_text(const ml_keys &k) {
    std::ostringstream os;
    os << '\n' << "-------------------\n" << "Multilevel Settings\n" << "-------------------\n";
    os << "coarse solver:   "
       << (k.cg == "LU" ? "cholesky\n" : k.cg == "redist" ? "redist\n" : "serial cedar\n");
    os << "relaxation:      " << k.relax_name << '\n';
    os << "cycle:           " << (k.cycle == CDR_CYCLE_V ? "V" : "F") << '\n';
    os << "min coarse:      " << k.min_coarse << '\n';
    if (k.num_levels > 0) os << "num levels:     " << k.num_levels << '\n';
    os << "nrelax pre:      " << k.nrelax_pre << '\n';
    os << "nrelax post:     " << k.nrelax_post << '\n';
    os << "maxiter:         " << k.maxiter << '\n';
    os << "tol:             " << k.tol;
    return os.str();
}

// Periodic images in the ghost layer, which Cedar's distributed objects get from their halo
// exchange: an operator's images as examples/basic-2d-ser/periodic.cc:56-78 writes them
// (interior rows for x, interior columns for y, every plane), a grid function's for every ghost
// row and then every ghost column.
void operator_images(double *so, int planes, std::size_t ii, std::size_t jj, int mask) {
    const std::size_t P = ii * jj, nx = ii - 2, ny = jj - 2;
    for (int k = 0; k < planes; ++k) {
        double *a = so + (std::size_t)k * P;
        if (mask & 1)
            for (std::size_t j = 1; j <= ny; ++j) {
                a[j * ii] = a[j * ii + nx];
                a[j * ii + nx + 1] = a[j * ii + 1];
            }
        if (mask & 2)
            for (std::size_t i = 1; i <= nx; ++i) {
                a[i] = a[ny * ii + i];
                a[(ny + 1) * ii + i] = a[ii + i];
            }
    }
}
void vector_images(double *q, std::size_t ii, std::size_t jj, int mask) {
    const std::size_t nx = ii - 2, ny = jj - 2;
    if (mask & 2)
        for (std::size_t i = 0; i < ii; ++i) {
            q[i] = q[ny * ii + i];
            q[(ny + 1) * ii + i] = q[ii + i];
        }
    if (mask & 1)
        for (std::size_t j = 0; j < jj; ++j) {
            q[j * ii] = q[j * ii + nx];
            q[j * ii + nx + 1] = q[j * ii + 1];
        }
}

// 3D: the images of every periodic direction over the full extents, y, x, then z (what a
// periodic halo exchange on one process leaves in the ghost layer)
void images3(double *a, std::size_t ii, std::size_t jj, std::size_t kk, int mask) {
    auto at = [&](std::size_t i, std::size_t j, std::size_t k) -> double & {
        return a[i + ii * (j + jj * k)];
    };
    if (mask & 2)
        for (std::size_t k = 0; k < kk; ++k)
            for (std::size_t i = 0; i < ii; ++i) {
                at(i, 0, k) = at(i, jj - 2, k);
                at(i, jj - 1, k) = at(i, 1, k);
            }
    if (mask & 1)
        for (std::size_t k = 0; k < kk; ++k)
            for (std::size_t j = 0; j < jj; ++j) {
                at(0, j, k) = at(ii - 2, j, k);
                at(ii - 1, j, k) = at(1, j, k);
            }
    if (mask & 4)
        for (std::size_t j = 0; j < jj; ++j)
            for (std::size_t i = 0; i < ii; ++i) {
                at(i, j, 0) = at(i, j, kk - 2);
                at(i, j, kk - 1) = at(i, j, 1);
            }
}
// the periodic images of an operator (every plane) or a grid function of either dimension
void fill_operator_images(double *so, int planes, const std::size_t len[3], int nd, int mask) {
    if (nd == 2) {
        operator_images(so, planes, len[0], len[1], mask);
        return;
    }
    const std::size_t P = len[0] * len[1] * len[2];
    for (int d = 0; d < planes; ++d) images3(so + (std::size_t)d * P, len[0], len[1], len[2], mask);
}
void fill_vector_images(double *q, const std::size_t len[3], int nd, int mask) {
    if (nd == 2)
        vector_images(q, len[0], len[1], mask);
    else
        images3(q, len[0], len[1], len[2], mask);
}

bool upload(double *dst, const double *src, std::size_t n) {
    return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
}
bool download(double *dst, const double *src, std::size_t n) {
    return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------------------------------
// config and logging
// ------------------------------------------------------------------------------------------
int cedar_config_create(const char *fname, cedar_config *newconfig) {
    if (!newconfig) return CEDAR_ERR_OTHER;
    *newconfig = CEDAR_CONFIG_NULL;
    if (!fname) return CEDAR_ERR_FNAME;
    std::ifstream f(fname, std::ios::binary);
    if (!f.good()) return CEDAR_ERR_FNAME;
    std::stringstream ss;
    ss << f.rdbuf();
    auto c = std::make_shared<config_obj>();
    std::string text = ss.str();
    parser p(text);
    if (!p.parse(c->root)) return CEDAR_ERR_CONFIG;
    *newconfig = create(KCONFIG, c);
    return CEDAR_SUCCESS;
}

int cedar_config_get_int(cedar_config conf, const char *confpath, int *val) {
    auto c = get<config_obj>(conf, KCONFIG);
    if (!c) return CEDAR_ERR_CONFIG;
    if (!confpath || !val) return CEDAR_ERR_CONFIG_PATH;
    int v;
    if (!as_int(c->root.at(confpath), v)) return CEDAR_ERR_CONFIG_PATH;
    *val = v;
    return CEDAR_SUCCESS;
}

int cedar_config_free(cedar_config *conf) {
    if (!conf) return CEDAR_ERR_CONFIG;
    return release(conf, KCONFIG, CEDAR_ERR_CONFIG, CEDAR_CONFIG_NULL);
}

int cedar_log_init(cedar_config conf) {
    auto c = get<config_obj>(conf, KCONFIG);
    if (!c) return CEDAR_ERR_CONFIG;
    static const std::map<std::string, unsigned> names = {
        {"status", LOG_STATUS}, {"info", LOG_INFO},   {"error", LOG_ERROR},
        {"memory", LOG_MEMORY}, {"debug", LOG_DEBUG}, {"timer", LOG_TIMER}};
    unsigned lvl = 0;
    const jvalue *l = c->root.at("log");
    if (l && l->kind == jvalue::ARR)
        for (auto &e : l->arr)
            if (e.kind == jvalue::STR) {
                auto it = names.find(e.s);
                if (it != names.end()) lvl |= it->second;
            }
    g_log_level = lvl;
    return CEDAR_SUCCESS;
}

// ------------------------------------------------------------------------------------------
// topologies
// ------------------------------------------------------------------------------------------
int cedar_topo_create2d(void *comm, unsigned int ngx, unsigned int ngy, unsigned int lnx[],
                        unsigned int lny[], int nprocx, int nprocy, cedar_topo *newtopo) {
    (void)comm;
    if (!newtopo) return CEDAR_ERR_TOPO;
    *newtopo = CEDAR_TOPO_NULL;
    if (nprocx != 1 || nprocy != 1 || !lnx || !lny || lnx[0] != ngx || lny[0] != ngy)
        return CEDAR_ERR_TOPO;
    auto t = std::make_shared<topo_obj>();
    t->nd = 2;
    t->n[0] = ngx;
    t->n[1] = ngy;
    t->n[2] = 1;
    if (!valid_topo_dims(t->n, 2)) return CEDAR_ERR_TOPO;
    *newtopo = create(KTOPO, t);
    return CEDAR_SUCCESS;
}

int cedar_topo_create3d(void *comm, unsigned int ngx, unsigned int ngy, unsigned int ngz,
                        unsigned int lnx[], unsigned int lny[], unsigned int lnz[], int nprocx,
                        int nprocy, int nprocz, cedar_topo *newtopo) {
    (void)comm;
    if (!newtopo) return CEDAR_ERR_TOPO;
    *newtopo = CEDAR_TOPO_NULL;
    if (nprocx != 1 || nprocy != 1 || nprocz != 1 || !lnx || !lny || !lnz || lnx[0] != ngx ||
        lny[0] != ngy || lnz[0] != ngz)
        return CEDAR_ERR_TOPO;
    auto t = std::make_shared<topo_obj>();
    t->nd = 3;
    t->n[0] = ngx;
    t->n[1] = ngy;
    t->n[2] = ngz;
    if (!valid_topo_dims(t->n, 3)) return CEDAR_ERR_TOPO;
    *newtopo = create(KTOPO, t);
    return CEDAR_SUCCESS;
}

int cedar_topo_free(cedar_topo *topo) {
    if (!topo) return CEDAR_ERR_TOPO;
    return release(topo, KTOPO, CEDAR_ERR_TOPO, CEDAR_TOPO_NULL);
}

// ------------------------------------------------------------------------------------------
// vectors
// ------------------------------------------------------------------------------------------
static int vec_create(cedar_topo topo, cedar_vec *vec, int nd) {
    if (!vec) return CEDAR_ERR_VEC;
    *vec = CEDAR_VEC_NULL;
    auto t = get<topo_obj>(topo, KTOPO);
    if (!t) return CEDAR_ERR_TOPO;
    if (t->nd != nd) return CEDAR_ERR_DIM;
    auto v = std::make_shared<vec_obj>();
    v->nd = nd;
    v->topo = topo;
    v->grid = t;
    std::size_t n = 1;
    for (int d = 0; d < 3; ++d) {
        v->len[d] = d < nd ? (std::size_t)t->n[d] + 2 : 1;
        n *= v->len[d];
    }
    v->data.assign(n, 0.0);
    incref(topo);
    *vec = create(KVEC, v);
    return CEDAR_SUCCESS;
}

int cedar_vec_create2d(cedar_topo topo, cedar_vec *vec) { return vec_create(topo, vec, 2); }
int cedar_vec_create3d(cedar_topo topo, cedar_vec *vec) { return vec_create(topo, vec, 3); }

int cedar_vec_baseptr(cedar_vec vec, cedar_real **base) {
    auto v = get<vec_obj>(vec, KVEC);
    if (!v) return CEDAR_ERR_VEC;
    if (!base) return CEDAR_ERR_OTHER;
    *base = v->data.data();
    return CEDAR_SUCCESS;
}

int cedar_vec_len2d(cedar_vec vec, cedar_len *ilen, cedar_len *jlen) {
    auto v = get<vec_obj>(vec, KVEC);
    if (!v) return CEDAR_ERR_VEC;
    if (v->nd != 2) return CEDAR_ERR_DIM;
    if (!ilen || !jlen) return CEDAR_ERR_OTHER;
    *ilen = (cedar_len)v->len[0];
    *jlen = (cedar_len)v->len[1];
    return CEDAR_SUCCESS;
}

int cedar_vec_len3d(cedar_vec vec, cedar_len *ilen, cedar_len *jlen, cedar_len *klen) {
    auto v = get<vec_obj>(vec, KVEC);
    if (!v) return CEDAR_ERR_VEC;
    if (v->nd != 3) return CEDAR_ERR_DIM;
    if (!ilen || !jlen || !klen) return CEDAR_ERR_OTHER;
    *ilen = (cedar_len)v->len[0];
    *jlen = (cedar_len)v->len[1];
    *klen = (cedar_len)v->len[2];
    return CEDAR_SUCCESS;
}

int cedar_vec_getdim(cedar_vec vec, int *nd) {
    auto v = get<vec_obj>(vec, KVEC);
    if (!v) return CEDAR_ERR_VEC;
    if (!nd) return CEDAR_ERR_OTHER;
    *nd = v->nd;
    return CEDAR_SUCCESS;
}

int cedar_vec_free(cedar_vec *vec) {
    if (!vec) return CEDAR_ERR_VEC;
    auto v = get<vec_obj>(*vec, KVEC);
    if (!v) return CEDAR_ERR_VEC;
    int topo = v->topo;
    int st = release(vec, KVEC, CEDAR_ERR_VEC, CEDAR_VEC_NULL);
    decref(topo);
    return st;
}

// ------------------------------------------------------------------------------------------
// matrices
// ------------------------------------------------------------------------------------------
static int mat_create(cedar_config conf, cedar_topo topo, int nd, int nstncl, cedar_mat *mat) {
    if (!mat) return CEDAR_ERR_MAT;
    *mat = CEDAR_MAT_NULL;
    auto t = get<topo_obj>(topo, KTOPO);
    if (!t) return CEDAR_ERR_TOPO;
    auto c = get<config_obj>(conf, KCONFIG);
    if (!c) return CEDAR_ERR_CONFIG;
    if (t->nd != nd) return CEDAR_ERR_DIM;
    auto m = std::make_shared<mat_obj>();
    m->nd = nd;
    m->topo = topo;
    m->conf = conf;
    m->grid = t;
    m->cfg = c;
    m->nstncl = nstncl;
    m->npts = 1;
    for (int d = 0; d < 3; ++d) {
        m->len[d] = d < nd ? (std::size_t)t->n[d] + 2 : 1;
        m->npts *= m->len[d];
    }
    m->so.assign((std::size_t)nstncl * m->npts, 0.0);
    incref(topo);
    incref(conf);
    *mat = create(KMAT, m);
    return CEDAR_SUCCESS;
}

int cedar_mat_create2d(cedar_config conf, cedar_topo topo, cedar_stencil_2d sten_type,
                       cedar_mat *mat) {
    if (sten_type != CEDAR_STENCIL_FIVE_PT && sten_type != CEDAR_STENCIL_NINE_PT) {
        if (mat) *mat = CEDAR_MAT_NULL;
        return CEDAR_ERR_STENCIL;
    }
    return mat_create(conf, topo, 2, sten_type == CEDAR_STENCIL_FIVE_PT ? 3 : 5, mat);
}

int cedar_mat_create3d(cedar_config conf, cedar_topo topo, cedar_stencil_3d sten_type,
                       cedar_mat *mat) {
    if (sten_type != CEDAR_STENCIL_SEVEN_PT && sten_type != CEDAR_STENCIL_XXVII_PT) {
        if (mat) *mat = CEDAR_MAT_NULL;
        return CEDAR_ERR_STENCIL;
    }
    return mat_create(conf, topo, 3, sten_type == CEDAR_STENCIL_SEVEN_PT ? 4 : 14, mat);
}

// vertex-based entries to symmetric storage: src/2d/interface/c/mat.cc:60-125
int cedar_mat_set2d(cedar_mat mat, unsigned int nvals, cedar_coord_2d coords[],
                    cedar_real vals[]) {
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    if (m->nd != 2) return CEDAR_ERR_DIM;
    if (nvals && (!coords || !vals)) return CEDAR_ERR_OTHER;
    const std::size_t ii = m->len[0];
    const bool five = m->nstncl == 3;
    m->dirty = true;
    for (unsigned n = 0; n < nvals; ++n) {
        if (coords[n].i >= m->grid->n[0] || coords[n].j >= m->grid->n[1]) return CEDAR_ERR_DIM;
        std::size_t ci = (std::size_t)coords[n].i + 1, cj = (std::size_t)coords[n].j + 1;
        int dir = (int)coords[n].dir;
        double v = dir != BMG2_C ? -vals[n] : vals[n];
        int plane;
        switch (dir) {
        case BMG2_C: plane = 0; break;
        case BMG2_W: plane = 1; break;
        case BMG2_S: plane = 2; break;
        case BMG2_E: plane = 1; ci++; break;
        case BMG2_N: plane = 2; cj++; break;
        case BMG2_SW: plane = 3; break;
        case BMG2_NW: plane = 4; cj++; break;
        case BMG2_NE: plane = 3; ci++; cj++; break;
        case BMG2_SE: plane = 4; ci++; break;
        default: return CEDAR_ERR_STENCIL;
        }
        if (five && plane > 2) return CEDAR_ERR_STENCIL;
        m->so[(std::size_t)plane * m->npts + cj * ii + ci] = v;
    }
    return CEDAR_SUCCESS;
}

int cedar_mat_set3d(cedar_mat mat, unsigned int nvals, cedar_coord_3d coords[],
                    cedar_real vals[]) {
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    if (m->nd != 3) return CEDAR_ERR_DIM;
    if (nvals && (!coords || !vals)) return CEDAR_ERR_OTHER;
    const std::size_t ii = m->len[0], jj = m->len[1];
    m->dirty = true;
    for (unsigned n = 0; n < nvals; ++n) {
        if (coords[n].i >= m->grid->n[0] || coords[n].j >= m->grid->n[1] ||
            coords[n].k >= m->grid->n[2])
            return CEDAR_ERR_DIM;
        int dir = (int)coords[n].dir;
        if (dir < 0 || dir >= m->nstncl) return CEDAR_ERR_STENCIL;
        std::size_t ci = (std::size_t)coords[n].i + 1, cj = (std::size_t)coords[n].j + 1,
                    ck = (std::size_t)coords[n].k + 1;
        m->so[(std::size_t)dir * m->npts + (ck * jj + cj) * ii + ci] =
            dir != BMG3_P ? -vals[n] : vals[n];
    }
    return CEDAR_SUCCESS;
}

int cedar_matvec(cedar_mat mat, cedar_vec x, cedar_vec y) {
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    auto xv = get<vec_obj>(x, KVEC);
    auto yv = get<vec_obj>(y, KVEC);
    if (!xv || !yv) return CEDAR_ERR_VEC;
    if (m->nd != xv->nd || m->nd != yv->nd) return CEDAR_ERR_DIM;
    if (xv->data.size() != m->npts || yv->data.size() != m->npts) return CEDAR_ERR_DIM;
    const std::size_t n = m->npts;
    if (!m->dso.reserve(m->so.size()) || !m->dx.reserve(n) || !m->dy.reserve(n))
        return CEDAR_ERR_OTHER;
    // the reference exchanges the halo of x first (src/2d/interface/c/mat.cc:172-183); on one
    // grid that fills the periodic images
    const int pmask = periodic_mask(m->cfg->root) & (m->nd == 2 ? 3 : 7);
    if (m->dirty) {
        const double *src = m->so.data();
        std::vector<double> tmp;
        if (pmask) {
            tmp = m->so;
            fill_operator_images(tmp.data(), m->nstncl, m->len, m->nd, pmask);
            src = tmp.data();
        }
        if (!upload(m->dso.p, src, m->so.size())) return CEDAR_ERR_OTHER;
        m->dirty = false;
    }
    const double *xs = xv->data.data();
    std::vector<double> xtmp;
    if (pmask) {
        xtmp = xv->data;
        fill_vector_images(xtmp.data(), m->len, m->nd, pmask);
        xs = xtmp.data();
    }
    if (!upload(m->dx.p, xs, n) || !upload(m->dy.p, yv->data.data(), n))
        return CEDAR_ERR_OTHER;
    int st = m->nd == 2
                 ? cdr_hip_matvec2d_f64(m->dso.p, m->nstncl, m->dx.p, m->dy.p, (int)m->len[0],
                                        (int)m->len[1], nullptr)
                 : cdr_hip_matvec3d_f64(m->dso.p, m->nstncl, m->dx.p, m->dy.p, (int)m->len[0],
                                        (int)m->len[1], (int)m->len[2], nullptr);
    if (st != CDR_OK) return CEDAR_ERR_OTHER;
    if (!download(yv->data.data(), m->dy.p, n)) return CEDAR_ERR_OTHER;
    return CEDAR_SUCCESS;
}

int cedar_mat_gettopo(cedar_mat mat, cedar_topo *topo) {
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) {
        if (topo) *topo = CEDAR_TOPO_NULL;
        return CEDAR_ERR_MAT;
    }
    if (!topo) return CEDAR_ERR_OTHER;
    *topo = m->topo;
    return CEDAR_SUCCESS;
}

int cedar_mat_dump(cedar_mat mat) {
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    if (m->nd == 2) {
        std::ofstream f("op-0-0.txt", std::ios::out | std::ios::trunc | std::ios::binary);
        if (!f) return CEDAR_ERR_OTHER;
        f << cedar::hip::stencil_text(m->so.data(), m->nstncl, (int)m->len[0], (int)m->len[1]);
        return f.good() ? CEDAR_SUCCESS : CEDAR_ERR_OTHER;
    }
    std::ofstream f("so-0.0.0");
    if (!f) return CEDAR_ERR_OTHER;
    f << m->nstncl << '\n' << m->so.size() << '\n';
    for (double v : m->so) f << v << '\n';
    return f.good() ? CEDAR_SUCCESS : CEDAR_ERR_OTHER;
}

int cedar_mat_free(cedar_mat *mat) {
    if (!mat) return CEDAR_ERR_MAT;
    auto m = get<mat_obj>(*mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    int topo = m->topo, conf = m->conf;
    int st = release(mat, KMAT, CEDAR_ERR_MAT, CEDAR_MAT_NULL);
    decref(topo);
    decref(conf);
    return st;
}

// ------------------------------------------------------------------------------------------
// solvers
// ------------------------------------------------------------------------------------------
int cedar_solver_create(cedar_mat mat, cedar_solver *solver) {
    if (!solver) return CEDAR_ERR_SOLVER;
    *solver = CEDAR_SOLVER_NULL;
    auto m = get<mat_obj>(mat, KMAT);
    if (!m) return CEDAR_ERR_MAT;
    if (!m->cfg) return CEDAR_ERR_CONFIG;
    const jvalue &root = m->cfg->root;
    std::string err;
    ml_keys k;
    int prec = CDR_PREC_F64;
    if (!read_ml(root, k, err) || !read_precision(root, prec, err)) {
        log_line(LOG_ERROR, err);
        return CEDAR_ERR_CONFIG;
    }
    const int pmask = periodic_mask(root);
    if (m->nd == 2 && (pmask & 4)) {
        log_line(LOG_ERROR, "grid.periodic: a 2D problem has no z direction");
        return CEDAR_ERR_CONFIG;
    }
    if (pmask != 0 && m->nd == 3 && k.relaxation != CDR_RELAX_POINT) {
        log_line(LOG_ERROR, "grid.periodic: 3D periodic problems take point relaxation");
        return CEDAR_ERR_CONFIG;
    }
    if (get_or(root, "gpu.autotune", false, as_bool)) cdr_hip_autotune(1);
    log_line(LOG_STATUS, settings_text(k));

    auto s = std::make_shared<solver_obj>();
    s->nd = m->nd;
    s->maxiter = k.maxiter;
    s->npts = m->npts;
    for (int d = 0; d < 3; ++d) s->len[d] = m->len[d];
    if (!s->dso.reserve(m->so.size()) || !s->dx.reserve(s->npts) || !s->db.reserve(s->npts))
        return CEDAR_ERR_OTHER;
    s->pmask = pmask;
    {
        std::vector<double> op = m->so;
        if (pmask) fill_operator_images(op.data(), m->nstncl, m->len, m->nd, pmask);
        if (!upload(s->dso.p, op.data(), op.size())) return CEDAR_ERR_OTHER;
    }
    int st;
    if (m->nd == 2) {
        if (k.relaxation >= CDR_RELAX_PLANE_XY) {
            log_line(LOG_ERROR, "Invalid relaxation");
            return CEDAR_ERR_CONFIG;
        }
        cdr_solver_settings cs;
        to_settings(k, root, cs);
        cs.precision = prec;
        cs.periodic = pmask;
        st = cdr_hip_solver2d_create(s->dso.p, m->nstncl, (int)m->grid->n[0],
                                     (int)m->grid->n[1], &cs, nullptr, &s->s2);
    } else {
        if (k.relaxation != CDR_RELAX_POINT && k.relaxation < CDR_RELAX_PLANE_XY) {
            log_line(LOG_ERROR, "Invalid relaxation");
            return CEDAR_ERR_CONFIG;
        }
        cdr_solver3_settings cs;
        cdr_hip_default_settings3(&cs);
        cs.relaxation = k.relaxation;
        cs.nrelax_pre = k.nrelax_pre;
        cs.nrelax_post = k.nrelax_post;
        cs.num_levels = k.num_levels;
        cs.maxiter = k.maxiter;
        cs.tol = k.tol;
        cs.min_coarse = k.min_coarse;
        cs.cycle = k.cycle;
        cs.precision = prec;
        cs.stall_cycles = std::max(0, get_or(root, "gpu.stall-cycles", 0, as_int));
        cs.periodic = pmask;
        // plane-config: its own ml settings when present, else line-xy with one cycle
        // (src/kernel_params.cc:79-85)
        const jvalue *pc = root.member("plane-config");
        if (pc && pc->kind == jvalue::OBJ && !pc->obj.empty()) {
            ml_keys pk;
            if (!read_ml(*pc, pk, err)) {
                log_line(LOG_ERROR, err);
                return CEDAR_ERR_CONFIG;
            }
            if (pk.relaxation >= CDR_RELAX_PLANE_XY) {
                log_line(LOG_ERROR, "Invalid relaxation");
                return CEDAR_ERR_CONFIG;
            }
            to_settings(pk, *pc, cs.plane);
        }
        st = cdr_hip_solver3d_create(s->dso.p, m->nstncl, (int)m->grid->n[0],
                                     (int)m->grid->n[1], (int)m->grid->n[2], &cs, nullptr,
                                     &s->s3);
    }
    if (st == CDR_ERR_LEVELS) {
        log_line(LOG_ERROR, "too many levels specified");
        return CEDAR_ERR_CONFIG;
    }
    if (st == CDR_ERR_ARG) return CEDAR_ERR_CONFIG;
    if (st != CDR_OK) return CEDAR_ERR_OTHER;
    *solver = create(KSOLVER, s);
    return CEDAR_SUCCESS;
}

int cedar_solver_run(cedar_solver solver, cedar_vec x, cedar_vec b) {
    auto s = get<solver_obj>(solver, KSOLVER);
    if (!s) return CEDAR_ERR_SOLVER;
    auto xv = get<vec_obj>(x, KVEC);
    auto bv = get<vec_obj>(b, KVEC);
    if (!xv || !bv) return CEDAR_ERR_VEC;
    if (s->nd != xv->nd || s->nd != bv->nd) return CEDAR_ERR_DIM;
    if (xv->data.size() != s->npts || bv->data.size() != s->npts) return CEDAR_ERR_DIM;
    const double *xs = xv->data.data(), *bs = bv->data.data();
    std::vector<double> xtmp, btmp;
    if (s->pmask) {  // images of x and b, as the reference's halo exchange provides them
        xtmp = xv->data;
        btmp = bv->data;
        fill_vector_images(xtmp.data(), s->len, s->nd, s->pmask);
        fill_vector_images(btmp.data(), s->len, s->nd, s->pmask);
        xs = xtmp.data();
        bs = btmp.data();
    }
    if (!upload(s->dx.p, xs, s->npts) || !upload(s->db.p, bs, s->npts)) return CEDAR_ERR_OTHER;
    std::vector<double> hist((std::size_t)(s->maxiter > 0 ? s->maxiter : 1));
    int nit = 0;
    int st = s->nd == 2 ? cdr_hip_solver2d_solve(s->s2, s->db.p, s->dx.p, hist.data(),
                                                 (int)hist.size(), &nit)
                        : cdr_hip_solver3d_solve(s->s3, s->db.p, s->dx.p, hist.data(),
                                                 (int)hist.size(), &nit);
    // like the reference, a solve whose residual turns non-finite still returns success; so
    // does one stopped by the stall guard (gpu.stall-cycles), which the reference lacks
    if (st != CDR_OK && st != CDR_ERR_NONFINITE && st != CDR_ERR_STALLED) return CEDAR_ERR_OTHER;
    if (!download(xv->data.data(), s->dx.p, s->npts)) return CEDAR_ERR_OTHER;
    for (int i = 0; i < nit; ++i) {
        std::ostringstream os;
        os << "Iteration " << i << " relative l2 norm: " << hist[(std::size_t)i];
        log_line(LOG_STATUS, os.str());
    }
    if (st == CDR_ERR_NONFINITE) log_line(LOG_ERROR, "non-finite residual norm");
    if (st == CDR_ERR_STALLED) log_line(LOG_ERROR, "residual stopped decreasing");
    return CEDAR_SUCCESS;
}

int cedar_solver_destroy(cedar_solver *solver) {
    if (!solver) return CEDAR_ERR_SOLVER;
    return release(solver, KSOLVER, CEDAR_ERR_SOLVER, CEDAR_SOLVER_NULL);
}

int cedar_solver_free(cedar_solver *solver) { return cedar_solver_destroy(solver); }

}  // extern "C"
