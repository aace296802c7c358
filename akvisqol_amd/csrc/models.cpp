// MOS mapping models on the host: the libsvm nu-SVR (svr.zig) and the
// TF-Lattice speech model (lattice.zig) through a small TFLite flatbuffer
// interpreter for the 18 builtin ops the model uses (tflite.zig).
//
// The lattice is one graph walk per pair (about 0.25 M f32 multiply-adds):
// cheaper on the host than a device round trip, so it stays here.  Dots
// are sequential f32 fused multiply-adds, the TFLite builtin order the
// reference reproduces (tflite.zig:665-671).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <cstdlib>
#include <sstream>

#include "host.h"

namespace vq {

// ---------------------------------------------------------------- SVR
static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t')) b--;
    return s.substr(a, b - a);
}

static bool parse_double(const std::string& s, double& v) {
    if (s.empty()) return false;
    char* end = nullptr;
    v = strtod(s.c_str(), &end);
    return end && *end == 0;
}

std::unique_ptr<SvrModel> SvrModel::parse(const std::string& text, std::string& err) {
    auto m = std::unique_ptr<SvrModel>(new SvrModel());
    std::vector<std::vector<std::string>> rows;
    bool in_sv = false;
    size_t max_feat = 0;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        if (!in_sv) {
            if (line == "SV") { in_sv = true; continue; }
            const size_t sp = line.find(' ');
            if (sp == std::string::npos) continue;
            const std::string key = line.substr(0, sp), val = trim(line.substr(sp + 1));
            if (key == "svm_type" && val != "nu_svr" && val != "epsilon_svr") { err = "UnsupportedSvmType"; return nullptr; }
            if (key == "kernel_type" && val != "rbf") { err = "UnsupportedKernel"; return nullptr; }
            if (key == "gamma" && !parse_double(val, m->gamma_)) { err = "InvalidModelFile"; return nullptr; }
            if (key == "rho" && !parse_double(val, m->rho_)) { err = "InvalidModelFile"; return nullptr; }
            continue;
        }
        std::vector<std::string> toks;
        std::string tok;
        std::istringstream ls(line);
        while (ls >> tok) toks.push_back(tok);
        for (size_t i = 1; i < toks.size(); i++) {
            const size_t colon = toks[i].find(':');
            if (colon == std::string::npos || colon == 0) continue;
            const std::string is = toks[i].substr(0, colon);
            if (is.find_first_not_of("0123456789") != std::string::npos) continue;
            max_feat = std::max(max_feat, (size_t)strtoull(is.c_str(), nullptr, 10));
        }
        rows.push_back(std::move(toks));
    }
    if (rows.empty()) { err = "NoSupportVectors"; return nullptr; }
    if (max_feat == 0) { err = "InvalidModelFile"; return nullptr; }
    m->nf_ = (int)max_feat;
    m->sv_.assign(rows.size() * max_feat, 0.0);
    m->coef_.resize(rows.size());
    for (size_t r = 0; r < rows.size(); r++) {
        if (rows[r].empty() || !parse_double(rows[r][0], m->coef_[r])) { err = "InvalidModelFile"; return nullptr; }
        for (size_t i = 1; i < rows[r].size(); i++) {
            const size_t colon = rows[r][i].find(':');
            if (colon == std::string::npos) continue;
            const std::string is = rows[r][i].substr(0, colon);
            if (is.empty() || is.find_first_not_of("0123456789") != std::string::npos) continue;
            double v;
            if (!parse_double(rows[r][i].substr(colon + 1), v)) continue;
            const size_t idx = strtoull(is.c_str(), nullptr, 10);
            if (idx == 0 || idx > max_feat) { err = "InvalidFeatureIndex"; return nullptr; }
            m->sv_[r * max_feat + idx - 1] = v;
        }
    }
    return m;
}

double SvrModel::predict(const double* x) const {
    double sum = 0.0;
    for (size_t i = 0; i < coef_.size(); i++) {
        const double* sv = &sv_[i * nf_];
        double dist = 0.0;
        for (int k = 0; k < nf_; k++) {
            const double d = sv[k] - x[k];
            dist += d * d;
        }
        sum += coef_[i] * exp(-gamma_ * dist);
    }
    return sum - rho_;
}

// ---------------------------------------------------------------- TFLite
namespace {

struct FbError {};

// Bounds-checked little-endian flatbuffer reader (tflite.zig:15-82).
struct Fb {
    const uint8_t* p;
    size_t n;
    template <typename T>
    T at(size_t pos) const {
        if (pos > n || n - pos < sizeof(T)) throw FbError{};
        T v;
        memcpy(&v, p + pos, sizeof(T));
        return v;
    }
    // Absolute position of a table field, 0 if absent.
    size_t field(size_t tab, int id) const {
        const long long vt = (long long)tab - at<int32_t>(tab);
        if (vt < 0) throw FbError{};
        const uint16_t vsize = at<uint16_t>((size_t)vt);
        const size_t slot = 4 + 2 * (size_t)id;
        if (slot + 2 > vsize) return 0;
        const uint16_t off = at<uint16_t>((size_t)vt + slot);
        return off ? tab + off : 0;
    }
    template <typename T>
    T scalar(size_t tab, int id, T dflt) const {
        const size_t f = field(tab, id);
        return f ? at<T>(f) : dflt;
    }
    size_t deref(size_t pos) const { return pos + at<uint32_t>(pos); }
    size_t table(size_t tab, int id) const {
        const size_t f = field(tab, id);
        return f ? deref(f) : 0;
    }
    // (start, len) of a vector field; len < 0 if absent.
    std::pair<size_t, long long> vec(size_t tab, int id) const {
        const size_t f = field(tab, id);
        if (!f) return {0, -1};
        const size_t v = deref(f);
        return {v + 4, (long long)at<uint32_t>(v)};
    }
    std::string str(size_t tab, int id) const {
        auto v = vec(tab, id);
        if (v.second < 0) return std::string();
        if (v.first > n || n - v.first < (size_t)v.second) throw FbError{};
        return std::string((const char*)p + v.first, (size_t)v.second);
    }
};

// Sequential fused dot product acc = fma(x[k], w[k], acc), k = 0..n-1: the
// same single-rounding operations whether fmaf is the FMA instruction or the
// C library's routine; the clone for FMA hardware keeps them inline.
__attribute__((target("fma"))) float dot_fma_hw(const float* x, const float* w, size_t n) {
    float acc = 0.0f;
    for (size_t k = 0; k < n; k++) acc = __builtin_fmaf(x[k], w[k], acc);
    return acc;
}
float dot_fma_lib(const float* x, const float* w, size_t n) {
    float acc = 0.0f;
    for (size_t k = 0; k < n; k++) acc = fmaf(x[k], w[k], acc);
    return acc;
}
float dot_fma(const float* x, const float* w, size_t n) {
    static const bool hw = __builtin_cpu_supports("fma");
    return hw ? dot_fma_hw(x, w, n) : dot_fma_lib(x, w, n);
}

// Eight such dot products side by side: each chain keeps its own order (the
// same bits as eight dot_fma calls), and the chains overlap in the FMA
// pipeline instead of each waiting on its own latency.
__attribute__((target("fma"))) void dot8_fma_hw(const float* const* x, const float* const* w, size_t n, float* out) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f, a5 = 0.0f, a6 = 0.0f, a7 = 0.0f;
    const float *x0 = x[0], *x1 = x[1], *x2 = x[2], *x3 = x[3], *x4 = x[4], *x5 = x[5], *x6 = x[6], *x7 = x[7];
    const float *w0 = w[0], *w1 = w[1], *w2 = w[2], *w3 = w[3], *w4 = w[4], *w5 = w[5], *w6 = w[6], *w7 = w[7];
    for (size_t k = 0; k < n; k++) {
        a0 = __builtin_fmaf(x0[k], w0[k], a0);
        a1 = __builtin_fmaf(x1[k], w1[k], a1);
        a2 = __builtin_fmaf(x2[k], w2[k], a2);
        a3 = __builtin_fmaf(x3[k], w3[k], a3);
        a4 = __builtin_fmaf(x4[k], w4[k], a4);
        a5 = __builtin_fmaf(x5[k], w5[k], a5);
        a6 = __builtin_fmaf(x6[k], w6[k], a6);
        a7 = __builtin_fmaf(x7[k], w7[k], a7);
    }
    out[0] = a0; out[1] = a1; out[2] = a2; out[3] = a3; out[4] = a4; out[5] = a5; out[6] = a6; out[7] = a7;
}
void dot8_fma(const float* const* x, const float* const* w, size_t n, float* out) {
    static const bool hw = __builtin_cpu_supports("fma");
    if (hw) {
        dot8_fma_hw(x, w, n, out);
        return;
    }
    for (int j = 0; j < 8; j++) out[j] = dot_fma_lib(x[j], w[j], n);
}

enum class Op { Add, Sub, Mul, Min, Max, Equal, Relu, Relu6, Abs, Cast, Concat, FullyConnected, BatchMatMul, Reshape,
                ExpandDims, Split, Shape, Fill, Unknown };
enum class Dt { None, F32, I32, Bool };

Op op_of(int code) {
    switch (code) {
        case 0: return Op::Add;
        case 2: return Op::Concat;
        case 9: return Op::FullyConnected;
        case 18: return Op::Mul;
        case 19: return Op::Relu;
        case 21: return Op::Relu6;
        case 22: return Op::Reshape;
        case 41: return Op::Sub;
        case 49: return Op::Split;
        case 53: return Op::Cast;
        case 55: return Op::Max;
        case 57: return Op::Min;
        case 70: return Op::ExpandDims;
        case 71: return Op::Equal;
        case 77: return Op::Shape;
        case 94: return Op::Fill;
        case 101: return Op::Abs;
        case 126: return Op::BatchMatMul;
        default: return Op::Unknown;
    }
}

// A tensor value: shape plus one of three typed payloads (f32 / i32 / bool
// stored as bytes).  Payloads are shared so reshape / expand are views.
struct Val {
    bool set = false;
    Dt dt = Dt::None;
    std::vector<size_t> shape;
    std::shared_ptr<std::vector<float>> f;
    std::shared_ptr<std::vector<int32_t>> i;
    std::shared_ptr<std::vector<uint8_t>> b;
    size_t count() const {
        size_t c = 1;
        for (size_t d : shape) c *= d;
        return c;
    }
};

// An elementwise node's broadcast for one pair of operand shapes: output
// shape and each output element's source indices.
struct BroadcastCache {
    bool ok = false;
    std::vector<size_t> xs, ys, shape, ix, iy;
};

struct Node {
    Op op;
    std::vector<int> in, out;
    int act = 0;
    int axis = 0;
    bool adj_x = false, adj_y = false, keep_dims = false;
    bool has_shape = false;
    std::vector<int32_t> new_shape;
    // filled by the model's warm-up call at load (single-threaded), read-only
    // afterwards; used only when the operands' shapes match
    mutable BroadcastCache bc;
};

inline float act_apply(int act, float x) {
    switch (act) {
        case 1: return fmaxf(x, 0.0f);
        case 2: return fmaxf(-1.0f, fminf(x, 1.0f));
        case 3: return fmaxf(0.0f, fminf(x, 6.0f));
        default: return x;
    }
}

struct GraphError {};

std::vector<size_t> bshape(const std::vector<size_t>& a, const std::vector<size_t>& b) {
    const size_t r = std::max(a.size(), b.size());
    std::vector<size_t> o(r);
    for (size_t k = 0; k < r; k++) {
        const size_t da = k + a.size() >= r ? a[k + a.size() - r] : 1;
        const size_t db = k + b.size() >= r ? b[k + b.size() - r] : 1;
        if (da != db && da != 1 && db != 1) throw GraphError{};
        o[k] = da == 1 ? db : da;
    }
    return o;
}

// Flat source index of every output element under numpy-style broadcast.
std::vector<size_t> bindex(const std::vector<size_t>& src, const std::vector<size_t>& out) {
    const size_t r = out.size();
    std::vector<size_t> stride(r, 0);
    size_t acc = 1;
    for (size_t k = src.size(); k-- > 0;) {
        const size_t ok = k + r - src.size();
        stride[ok] = src[k] == 1 ? 0 : acc;
        acc *= src[k];
    }
    size_t total = 1;
    for (size_t d : out) total *= d;
    std::vector<size_t> idx(total);
    std::vector<size_t> pos(r, 0);
    size_t cur = 0;
    for (size_t e = 0; e < total; e++) {
        idx[e] = cur;
        for (size_t k = r; k-- > 0;) {
            pos[k]++;
            cur += stride[k];
            if (pos[k] < out[k]) break;
            cur -= stride[k] * out[k];
            pos[k] = 0;
        }
    }
    return idx;
}

}  // namespace

class LatticeModel {
   public:
    std::vector<Dt> dtypes;
    std::vector<Val> consts;
    std::vector<Node> nodes;
    std::vector<std::pair<std::string, int>> inputs, outputs;
    std::vector<int> feat;
    int tau = -1, output = -1, nb = 0;

    int find(const std::vector<std::pair<std::string, int>>& l, const std::string& name) const {
        for (auto& e : l)
            if (e.first == name) return e.second;
        return -1;
    }

    // release[i]: tensors whose last reader is node i (never the output);
    // dropping them there keeps a call's working set small and its heap
    // blocks reused instead of returned to the system and faulted back in.
    std::vector<std::vector<int>> release;
    // Tensors a call sets that no node releases (never read, or the output):
    // cleared after the call so the per-thread workspace starts empty.
    std::vector<int> unreleased;
    bool warming = false;  // the load-time call that fills the broadcast caches

    void plan_release() {
        std::vector<int> last(consts.size(), -1);
        for (size_t i = 0; i < nodes.size(); i++)
            for (int t : nodes[i].in)
                if (t >= 0 && (size_t)t < last.size()) last[t] = (int)i;
        release.assign(nodes.size(), {});
        for (size_t t = 0; t < last.size(); t++)
            if (last[t] >= 0 && (int)t != output) release[last[t]].push_back((int)t);
        std::vector<char> set(consts.size(), 0);
        for (const Node& n : nodes)
            for (int t : n.out)
                if (t >= 0 && (size_t)t < set.size()) set[t] = 1;
        for (int t : feat) set[t] = 1;
        if (tau >= 0) set[tau] = 1;
        unreleased.clear();
        for (size_t t = 0; t < set.size(); t++)
            if (set[t] && (last[t] < 0 || (int)t == output)) unreleased.push_back((int)t);
    }

    // keep: leave every tensor's value in v (the recording call that the
    // execution plan is built from)
    void run(std::vector<Val>& v, bool keep = false) const {
        for (size_t i = 0; i < nodes.size(); i++) {
            exec(nodes[i], v);
            if (!keep)
                for (int t : release[i]) v[t] = Val{};
        }
    }

    // ---------------------------------------------------------------- execution plan
    // The graph's shapes do not depend on the features, so after a recording
    // call every tensor that does not depend on them (constants, shape
    // arithmetic, tau) is folded to its recorded value, and the rest runs
    // as a list of steps over one float arena: no allocation, no shape or
    // broadcast work per call, reshapes as aliases, slots reused once a
    // tensor's last reader has run.  Each step does the interpreter's
    // arithmetic on the same elements in the same order, so predictions are
    // bit-identical; a graph the plan cannot express (integer tensors that
    // depend on the features, say) keeps the interpreter.
    struct Operand {
        const float* p = nullptr;  // folded value, or
        int off = -1;              // arena offset
    };
    struct Step {
        Op op = Op::Unknown;
        int act = 0;
        int out = 0;
        size_t n = 0;  // output elements
        Operand x, y, bias;
        const size_t* ix = nullptr;  // broadcast maps (nullptr: dense, or scalar when the count is 1)
        const size_t* iy = nullptr;
        size_t xc = 0, yc = 0;       // operand element counts
        size_t m = 0, kk = 0, nn = 0, batch = 0, nout = 0, nin = 0;
        bool adj_x = false, adj_y = false;
        std::vector<Operand> ins;    // concat inputs
        std::vector<size_t> chunks;  // concat: per input chunk; split: chunk
        std::vector<int> outs;       // split outputs
        size_t outer = 1;
        // fused single-output fully-connected steps: operands per product
        std::vector<Operand> fx, fw, fb;
        std::vector<int> fout, fact;
    };
    std::vector<Step> plan;
    std::vector<std::vector<float>> plan_values;  // folded tensors the plan reads
    std::vector<std::vector<size_t>> plan_maps;
    std::vector<int> plan_feat;  // arena offsets of the feature inputs
    size_t plan_size = 0;
    int plan_out = -1;
    bool plan_ok = false;

    void build_plan(const std::vector<Val>& rec) {
        plan_ok = false;
        const size_t T = consts.size();
        std::vector<char> dyn(T, 0);
        for (int t : feat) dyn[t] = 1;
        for (const Node& n : nodes) {
            bool d = false;
            for (int t : n.in)
                if (t >= 0 && (size_t)t < T && dyn[t]) d = true;
            if (n.op == Op::Shape) d = false;
            for (int t : n.out)
                if (t >= 0 && (size_t)t < T) dyn[t] = d;
        }
        if (output < 0 || !dyn[output]) return;
        auto value = [&](int t) -> const Val& { return rec[t].set ? rec[t] : consts[t]; };
        // Execution order: the graph's, except that each long single-output
        // fully-connected node moves to just before its first reader (its
        // inputs are ready earlier, and nothing between reads it), so that
        // such nodes end up side by side and run together.
        const int N = (int)nodes.size();
        std::vector<int> first_reader(T, N);
        for (int i = N - 1; i >= 0; i--)
            for (int t : nodes[i].in)
                if (t >= 0 && (size_t)t < T) first_reader[t] = i;
        std::vector<std::pair<double, int>> keyed(N);
        for (int i = 0; i < N; i++) {
            const Node& n = nodes[i];
            double key = i;
            if (n.op == Op::FullyConnected && n.in.size() >= 2 && n.out.size() == 1 && n.out[0] >= 0 &&
                dyn[n.out[0]]) {
                const Val& w = value(n.in[1]);
                if (w.dt == Dt::F32 && w.shape.size() == 2 && w.shape[0] == 1 && w.shape[1] >= 256 &&
                    value(n.in[0]).count() == w.shape[1] && first_reader[n.out[0]] < N)
                    key = first_reader[n.out[0]] - 0.5;
            }
            keyed[i] = {key, i};
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
        std::vector<int> ord(N);
        for (int i = 0; i < N; i++) ord[i] = keyed[i].second;
        // storage: aliases share their source's slot; a slot is free after
        // the last reader of any tensor on it
        std::vector<int> owner(T, -1);
        std::vector<int> last(T, -1);
        for (int t : feat) owner[t] = t;
        for (size_t i = 0; i < nodes.size(); i++) {
            const Node& n = nodes[i];
            for (int t : n.out)
                if (t >= 0 && dyn[t]) owner[t] = t;
            if ((n.op == Op::Reshape || n.op == Op::ExpandDims) && dyn[n.in[0]] && !n.out.empty())
                owner[n.out[0]] = owner[n.in[0]];
        }
        for (int pos = 0; pos < N; pos++)
            for (int t : nodes[ord[pos]].in)
                if (t >= 0 && (size_t)t < T && dyn[t]) last[owner[t]] = std::max(last[owner[t]], pos);
        last[owner[output]] = N;  // never freed
        std::vector<int> off(T, -1);
        std::vector<char> freed(T, 0);
        std::vector<std::pair<size_t, size_t>> free_list;  // (offset, size)
        size_t top = 0;
        auto alloc = [&](size_t cnt) -> int {
            const size_t sz = (cnt + 15) & ~(size_t)15;
            for (auto it = free_list.begin(); it != free_list.end(); ++it)
                if (it->second >= sz) {
                    const size_t o = it->first;
                    it->first += sz;
                    it->second -= sz;
                    if (it->second == 0) free_list.erase(it);
                    return (int)o;
                }
            const size_t o = top;
            top += sz;
            return (int)o;
        };
        auto release_slot = [&](int t) {
            const size_t sz = (value(t).count() + 15) & ~(size_t)15;
            free_list.emplace_back((size_t)off[t], sz);
        };
        auto operand = [&](int t, size_t& count) -> Operand {
            const Val& v = value(t);
            count = v.count();
            Operand o;
            if (dyn[t]) {
                o.off = off[owner[t]];
                return o;
            }
            if (v.dt != Dt::F32 || !v.f) throw GraphError{};
            if (consts[t].set) {
                o.p = consts[t].f->data();  // the model's own constant
                return o;
            }
            plan_values.push_back(*v.f);
            o.p = plan_values.back().data();
            return o;
        };
        try {
            plan.clear();
            plan_values.clear();
            plan_maps.clear();
            plan_feat.clear();
            plan_values.reserve(nodes.size() * 3 + 4);  // pointers into it stay valid
            for (int t : feat) {
                if (value(t).count() != 1) return;
                off[t] = alloc(1);
                plan_feat.push_back(off[t]);
            }
            for (int i = 0; i < N; i++) {
                const Node& n = nodes[ord[i]];
                bool d = false;
                for (int t : n.out)
                    if (t >= 0 && dyn[t]) d = true;
                if (d) {
                    for (int t : n.out)
                        if (t >= 0 && owner[t] == t) {
                            if (value(t).dt != Dt::F32 && !(n.op == Op::Equal && value(t).dt == Dt::Bool)) return;
                            off[t] = alloc(value(t).count());
                        }
                    if (!plan_step(n, rec, dyn, operand, off, owner)) return;
                }
                for (int t : n.in)
                    if (t >= 0 && (size_t)t < T && dyn[t]) {
                        const int r = owner[t];
                        if (last[r] == (int)i && !freed[r]) {
                            release_slot(r);
                            freed[r] = 1;
                        }
                    }
                for (int t : n.out)  // written but never read
                    if (t >= 0 && dyn[t] && owner[t] == t && last[t] < 0 && !freed[t]) {
                        release_slot(t);
                        freed[t] = 1;
                    }
            }
        } catch (const GraphError&) {
            plan.clear();
            return;
        }
        // consecutive single-output products of one length become one step
        std::vector<Step> fused;
        for (size_t i = 0; i < plan.size(); i++) {
            Step& st = plan[i];
            const bool single = st.op == Op::FullyConnected && st.batch == 1 && st.nout == 1 && st.nin >= 256;
            auto reads_group = [&](const Step& g) {  // a product of the group feeds this one
                for (int o : g.fout)
                    if ((!st.x.p && st.x.off == o) || (!st.y.p && st.y.off == o) ||
                        (st.kk && !st.bias.p && st.bias.off == o))
                        return true;
                return false;
            };
            if (single && !fused.empty() && fused.back().op == Op::FullyConnected && !fused.back().fx.empty() &&
                fused.back().nin == st.nin && !reads_group(fused.back())) {
                Step& g = fused.back();
                g.fx.push_back(st.x);
                g.fw.push_back(st.y);
                g.fb.push_back(st.kk ? st.bias : Operand{});
                g.fout.push_back(st.out);
                g.fact.push_back(st.act);
                continue;
            }
            if (single) {
                Step g;
                g.op = Op::FullyConnected;
                g.nin = st.nin;
                g.fx.push_back(st.x);
                g.fw.push_back(st.y);
                g.fb.push_back(st.kk ? st.bias : Operand{});
                g.fout.push_back(st.out);
                g.fact.push_back(st.act);
                fused.push_back(std::move(g));
                continue;
            }
            fused.push_back(std::move(st));
        }
        plan = std::move(fused);
        plan_size = top;
        plan_out = off[owner[output]];
        plan_ok = true;
    }

    template <typename OperandFn>
    bool plan_step(const Node& n, const std::vector<Val>& rec, const std::vector<char>& dyn, OperandFn& operand,
                   const std::vector<int>& off, const std::vector<int>& owner) {
        auto value = [&](int t) -> const Val& { return rec[t].set ? rec[t] : consts[t]; };
        Step st;
        st.op = n.op;
        st.act = n.act;
        if (n.out.empty()) return false;
        const Val& outv = value(n.out[0]);
        st.out = off[owner[n.out[0]]];
        st.n = outv.count();
        auto map_of = [&](const Val& src) -> const size_t* {
            if (src.count() == st.n && src.shape == outv.shape) return nullptr;
            if (src.count() == 1) return nullptr;
            plan_maps.push_back(bindex(src.shape, outv.shape));
            return plan_maps.back().data();
        };
        switch (n.op) {
            case Op::Add: case Op::Sub: case Op::Mul: case Op::Min: case Op::Max: case Op::Equal: {
                const Val &x = value(n.in[0]), &y = value(n.in[1]);
                if (x.dt != Dt::F32 || y.dt != Dt::F32) return false;
                st.x = operand(n.in[0], st.xc);
                st.y = operand(n.in[1], st.yc);
                st.ix = map_of(x);
                st.iy = map_of(y);
                break;
            }
            case Op::Relu: case Op::Relu6: case Op::Abs: {
                if (value(n.in[0]).dt != Dt::F32) return false;
                st.x = operand(n.in[0], st.xc);
                break;
            }
            case Op::Cast: {
                const Val& x = value(n.in[0]);
                if (outv.dt != Dt::F32 || (x.dt != Dt::F32 && x.dt != Dt::Bool) || !dyn[n.in[0]]) return false;
                st.x = operand(n.in[0], st.xc);
                break;
            }
            case Op::Reshape: case Op::ExpandDims:
                return true;  // an alias of its input's slot: nothing to run
            case Op::Concat: {
                const int ax = norm_axis(n.axis, outv.shape.size());
                for (int d = 0; d < ax; d++) st.outer *= outv.shape[d];
                for (size_t k = 0; k < n.in.size(); k++) {
                    const Val& p = value(n.in[k]);
                    if (p.dt != Dt::F32) return false;
                    size_t inner = 1;
                    for (size_t d = ax + 1; d < p.shape.size(); d++) inner *= p.shape[d];
                    size_t c = 0;
                    st.ins.push_back(operand(n.in[k], c));
                    st.chunks.push_back(p.shape[ax] * inner);
                }
                break;
            }
            case Op::FullyConnected: {
                const Val &x = value(n.in[0]), &w = value(n.in[1]);
                if (x.dt != Dt::F32 || w.dt != Dt::F32 || w.shape.size() != 2) return false;
                st.nout = w.shape[0];
                st.nin = w.shape[1];
                st.batch = x.count() / st.nin;
                st.x = operand(n.in[0], st.xc);
                st.y = operand(n.in[1], st.yc);
                if (n.in.size() > 2 && n.in[2] >= 0) {
                    size_t c = 0;
                    st.bias = operand(n.in[2], c);
                    st.kk = 1;  // has bias
                }
                break;
            }
            case Op::BatchMatMul: {
                const Val &x = value(n.in[0]), &y = value(n.in[1]);
                if (x.dt != Dt::F32 || y.dt != Dt::F32) return false;
                const size_t xr = x.shape.size(), yr = y.shape.size();
                st.adj_x = n.adj_x;
                st.adj_y = n.adj_y;
                st.m = n.adj_x ? x.shape[xr - 1] : x.shape[xr - 2];
                st.kk = n.adj_x ? x.shape[xr - 2] : x.shape[xr - 1];
                st.nn = n.adj_y ? y.shape[yr - 2] : y.shape[yr - 1];
                std::vector<size_t> xb(x.shape.begin(), x.shape.end() - 2), yb(y.shape.begin(), y.shape.end() - 2);
                auto ob = bshape(xb, yb);
                std::vector<size_t> ix{0}, iy{0};
                if (!ob.empty()) {
                    ix = bindex(xb, ob);
                    iy = bindex(yb, ob);
                }
                st.batch = ix.size();
                plan_maps.push_back(std::move(ix));
                st.ix = plan_maps.back().data();
                plan_maps.push_back(std::move(iy));
                st.iy = plan_maps.back().data();
                st.x = operand(n.in[0], st.xc);
                st.y = operand(n.in[1], st.yc);
                break;
            }
            case Op::Split: {
                const Val &a = value(n.in[0]), &x = value(n.in[1]);
                if (a.dt != Dt::I32 || dyn[n.in[0]] || x.dt != Dt::F32) return false;
                const int ax = norm_axis((*a.i)[0], x.shape.size());
                const size_t parts = n.out.size();
                size_t inner = 1;
                for (int d = 0; d < ax; d++) st.outer *= x.shape[d];
                for (size_t d = ax + 1; d < x.shape.size(); d++) inner *= x.shape[d];
                st.chunks.push_back(x.shape[ax] / parts * inner);
                st.x = operand(n.in[1], st.xc);
                for (int t : n.out) st.outs.push_back(t >= 0 && owner[t] >= 0 ? off[owner[t]] : -1);
                break;
            }
            default:
                return false;
        }
        plan.push_back(std::move(st));
        return true;
    }

    // One elementwise step: z = f(x, y) per output element (the activation
    // after it unless `act` is false), with the operands' element maps
    // resolved once per step rather than per element.
    template <typename F>
    static void binary(const Step& st, const float* xf, const float* yf, float* o, F f, bool act = true) {
        const size_t n = st.n;
        const size_t* ix = st.ix;
        const size_t* iy = st.iy;
        const bool xs = !ix && st.xc == 1 && n != 1, ys = !iy && st.yc == 1 && n != 1;
        auto put = [&](size_t e, float z) { o[e] = act ? act_apply(st.act, z) : z; };
        if (!ix && !iy && !xs && !ys) {
            for (size_t e = 0; e < n; e++) put(e, f(xf[e], yf[e]));
        } else if (!ix && !xs && ys) {
            const float q = yf[0];
            for (size_t e = 0; e < n; e++) put(e, f(xf[e], q));
        } else if (xs && !iy && !ys) {
            const float p = xf[0];
            for (size_t e = 0; e < n; e++) put(e, f(p, yf[e]));
        } else {
            for (size_t e = 0; e < n; e++) {
                const float p = ix ? xf[ix[e]] : xs ? xf[0] : xf[e];
                const float q = iy ? yf[iy[e]] : ys ? yf[0] : yf[e];
                put(e, f(p, q));
            }
        }
    }

    void run_plan(float* A) const {
        auto get = [A](const Operand& o) -> const float* { return o.p ? o.p : A + o.off; };
        for (const Step& st : plan) {
            float* o = A + st.out;
            switch (st.op) {
                case Op::Add: case Op::Sub: case Op::Mul: case Op::Min: case Op::Max: case Op::Equal: {
                    const float* xf = get(st.x);
                    const float* yf = get(st.y);
                    switch (st.op) {
                        case Op::Add: binary(st, xf, yf, o, [](float p, float q) { return p + q; }); break;
                        case Op::Sub: binary(st, xf, yf, o, [](float p, float q) { return p - q; }); break;
                        case Op::Mul: binary(st, xf, yf, o, [](float p, float q) { return p * q; }); break;
                        case Op::Min: binary(st, xf, yf, o, [](float p, float q) { return fminf(p, q); }); break;
                        case Op::Max: binary(st, xf, yf, o, [](float p, float q) { return fmaxf(p, q); }); break;
                        default: binary(st, xf, yf, o, [](float p, float q) { return p == q ? 1.0f : 0.0f; }, false);
                    }
                    break;
                }
                case Op::Relu: case Op::Relu6: case Op::Abs: {
                    const float* xf = get(st.x);
                    for (size_t e = 0; e < st.n; e++) {
                        const float a = xf[e];
                        o[e] = st.op == Op::Relu ? fmaxf(a, 0.0f) : st.op == Op::Relu6 ? fmaxf(0.0f, fminf(a, 6.0f)) : fabsf(a);
                    }
                    break;
                }
                case Op::Cast: {
                    const float* xf = get(st.x);
                    for (size_t e = 0; e < st.n; e++) o[e] = xf[e];
                    break;
                }
                case Op::Concat: {
                    size_t w = 0;
                    for (size_t oi = 0; oi < st.outer; oi++)
                        for (size_t k = 0; k < st.ins.size(); k++) {
                            const float* src = get(st.ins[k]) + oi * st.chunks[k];
                            for (size_t e = 0; e < st.chunks[k]; e++) o[w++] = act_apply(st.act, src[e]);
                        }
                    break;
                }
                case Op::FullyConnected: {
                    if (!st.fx.empty()) {
                        const size_t k = st.fx.size();
                        size_t g = 0;
                        for (; g + 8 <= k; g += 8) {
                            const float* xs[8];
                            const float* ws[8];
                            float acc[8];
                            for (int j = 0; j < 8; j++) {
                                xs[j] = get(st.fx[g + j]);
                                ws[j] = get(st.fw[g + j]);
                            }
                            dot8_fma(xs, ws, st.nin, acc);
                            for (int j = 0; j < 8; j++) {
                                const Operand& b = st.fb[g + j];
                                float a = acc[j];
                                if (b.p || b.off >= 0) a += get(b)[0];
                                A[st.fout[g + j]] = act_apply(st.fact[g + j], a);
                            }
                        }
                        for (; g < k; g++) {
                            float a = dot_fma(get(st.fx[g]), get(st.fw[g]), st.nin);
                            const Operand& b = st.fb[g];
                            if (b.p || b.off >= 0) a += get(b)[0];
                            A[st.fout[g]] = act_apply(st.fact[g], a);
                        }
                        break;
                    }
                    const float* xf = get(st.x);
                    const float* wf = get(st.y);
                    const float* bf = st.kk ? get(st.bias) : nullptr;
                    for (size_t rr = 0; rr < st.batch; rr++)
                        for (size_t oo = 0; oo < st.nout; oo++) {
                            float acc = dot_fma(xf + rr * st.nin, wf + oo * st.nin, st.nin);
                            if (bf) acc += bf[oo];
                            o[rr * st.nout + oo] = act_apply(st.act, acc);
                        }
                    break;
                }
                case Op::BatchMatMul: {
                    const float* xf = get(st.x);
                    const float* yf = get(st.y);
                    const size_t m = st.m, kk = st.kk, nn = st.nn;
                    for (size_t bi = 0; bi < st.batch; bi++) {
                        const float* __restrict__ xm = xf + st.ix[bi] * m * kk;
                        const float* ym = yf + st.iy[bi] * kk * nn;
                        float* __restrict__ ob = o + bi * m * nn;
                        if (kk == 1) {
                            if (nn == 2) {
                                const float y0 = ym[0], y1 = ym[1];
                                for (size_t a = 0; a < m; a++) {
                                    ob[2 * a] = xm[a] * y0 + 0.0f;
                                    ob[2 * a + 1] = xm[a] * y1 + 0.0f;
                                }
                            } else {
                                for (size_t a = 0; a < m; a++)
                                    for (size_t c = 0; c < nn; c++) ob[a * nn + c] = xm[a] * ym[c] + 0.0f;
                            }
                            continue;
                        }
                        for (size_t a = 0; a < m; a++)
                            for (size_t c = 0; c < nn; c++) {
                                float acc = 0.0f;
                                for (size_t l = 0; l < kk; l++) {
                                    const float p = st.adj_x ? xm[l * m + a] : xm[a * kk + l];
                                    const float q = st.adj_y ? ym[c * kk + l] : ym[l * nn + c];
                                    acc = fmaf(p, q, acc);
                                }
                                ob[a * nn + c] = acc;
                            }
                    }
                    break;
                }
                case Op::Split: {
                    const float* xf = get(st.x);
                    const size_t parts = st.outs.size(), chunk = st.chunks[0];
                    for (size_t p = 0; p < parts; p++) {
                        if (st.outs[p] < 0) continue;
                        float* dst = A + st.outs[p];
                        for (size_t oi = 0; oi < st.outer; oi++)
                            memcpy(dst + oi * chunk, xf + (oi * parts + p) * chunk, sizeof(float) * chunk);
                    }
                    break;
                }
                default:
                    break;
            }
        }
    }

   private:
    // Tensor k of node n: the value computed in this call, else the model's
    // constant (the per-call table starts empty; constants are never copied).
    const Val& in(const std::vector<Val>& v, const Node& n, size_t k) const {
        if (k >= n.in.size() || n.in[k] < 0) throw GraphError{};
        const Val& a = v[n.in[k]];
        if (a.set) return a;
        const Val& c = consts[n.in[k]];
        if (!c.set) throw GraphError{};
        return c;
    }
    static Val floats(std::vector<size_t> shape, std::vector<float> data) {
        Val r;
        r.set = true;
        r.dt = Dt::F32;
        r.shape = std::move(shape);
        r.f = std::make_shared<std::vector<float>>(std::move(data));
        return r;
    }
    static int norm_axis(int axis, size_t rank) {
        const int a = axis < 0 ? axis + (int)rank : axis;
        if (a < 0 || a >= (int)rank) throw GraphError{};
        return a;
    }

    void exec(const Node& n, std::vector<Val>& v) const {
        if (n.out.empty()) throw GraphError{};
        Val r;
        switch (n.op) {
            case Op::Add: case Op::Sub: case Op::Mul: case Op::Min: case Op::Max: {
                const Val &x = in(v, n, 0), &y = in(v, n, 1);
                if (x.dt != Dt::F32 || y.dt != Dt::F32) throw GraphError{};
                BroadcastCache& bc = n.bc;
                const bool hit = bc.ok && x.shape == bc.xs && y.shape == bc.ys;
                auto shape = hit ? bc.shape : bshape(x.shape, y.shape);
                size_t total = 1;
                for (size_t d : shape) total *= d;
                std::vector<float> o(total);
                auto apply = [&](auto px, auto qy) {
                    for (size_t e = 0; e < total; e++) {
                        const float p = px(e), q = qy(e);
                        float z;
                        switch (n.op) {
                            case Op::Add: z = p + q; break;
                            case Op::Sub: z = p - q; break;
                            case Op::Mul: z = p * q; break;
                            case Op::Min: z = fminf(p, q); break;
                            default: z = fmaxf(p, q); break;
                        }
                        o[e] = act_apply(n.act, z);
                    }
                };
                const float* xf = x.f->data();
                const float* yf = y.f->data();
                const size_t xc = x.f->size(), yc = y.f->size();
                // same element order without index maps when an operand is
                // the whole output or a single value
                auto dense = [&](const float* f) { return [f](size_t e) { return f[e]; }; };
                auto scalar = [&](const float* f) { return [f](size_t) { return f[0]; }; };
                if (xc == total && yc == total && x.shape == y.shape) apply(dense(xf), dense(yf));
                else if (xc == total && x.shape == shape && yc == 1) apply(dense(xf), scalar(yf));
                else if (yc == total && y.shape == shape && xc == 1) apply(scalar(xf), dense(yf));
                else if (hit) {
                    const size_t* ix = bc.ix.data();
                    const size_t* iy = bc.iy.data();
                    apply([&](size_t e) { return xf[ix[e]]; }, [&](size_t e) { return yf[iy[e]]; });
                } else {
                    auto ix = bindex(x.shape, shape), iy = bindex(y.shape, shape);
                    apply([&](size_t e) { return xf[ix[e]]; }, [&](size_t e) { return yf[iy[e]]; });
                    if (warming) {
                        bc.ix = std::move(ix);
                        bc.iy = std::move(iy);
                    }
                }
                if (warming && !hit) {
                    bc.xs = x.shape;
                    bc.ys = y.shape;
                    bc.shape = shape;
                    bc.ok = true;
                }
                r = floats(std::move(shape), std::move(o));
                break;
            }
            case Op::Equal: {
                const Val &x = in(v, n, 0), &y = in(v, n, 1);
                if (x.dt != y.dt || x.dt == Dt::Bool) throw GraphError{};
                auto shape = bshape(x.shape, y.shape);
                auto ix = bindex(x.shape, shape), iy = bindex(y.shape, shape);
                auto o = std::make_shared<std::vector<uint8_t>>(ix.size());
                for (size_t e = 0; e < ix.size(); e++)
                    (*o)[e] = x.dt == Dt::F32 ? (*x.f)[ix[e]] == (*y.f)[iy[e]] : (*x.i)[ix[e]] == (*y.i)[iy[e]];
                r.set = true;
                r.dt = Dt::Bool;
                r.shape = shape;
                r.b = o;
                break;
            }
            case Op::Relu: case Op::Relu6: case Op::Abs: {
                const Val& x = in(v, n, 0);
                if (x.dt != Dt::F32) throw GraphError{};
                std::vector<float> o(x.f->size());
                for (size_t e = 0; e < o.size(); e++) {
                    const float a = (*x.f)[e];
                    o[e] = n.op == Op::Relu ? fmaxf(a, 0.0f) : n.op == Op::Relu6 ? fmaxf(0.0f, fminf(a, 6.0f)) : fabsf(a);
                }
                r = floats(x.shape, std::move(o));
                break;
            }
            case Op::Cast: {
                const Val& x = in(v, n, 0);
                const Dt to = dtypes[n.out[0]];
                const size_t cnt = x.count();
                r.set = true;
                r.dt = to;
                r.shape = x.shape;
                auto get = [&](size_t e) -> double {
                    if (x.dt == Dt::F32) return (*x.f)[e];
                    if (x.dt == Dt::I32) return (*x.i)[e];
                    return (*x.b)[e];
                };
                if (to == Dt::F32) {
                    r.f = std::make_shared<std::vector<float>>(cnt);
                    for (size_t e = 0; e < cnt; e++) (*r.f)[e] = x.dt == Dt::F32 ? (*x.f)[e] : (float)get(e);
                } else if (to == Dt::I32) {
                    r.i = std::make_shared<std::vector<int32_t>>(cnt);
                    for (size_t e = 0; e < cnt; e++) {
                        if (x.dt == Dt::I32) { (*r.i)[e] = (*x.i)[e]; continue; }
                        const double t = trunc(get(e));  // toward zero, saturating (tflite.zig:595-596)
                        (*r.i)[e] = isnan(t) ? 0 : t >= 2147483647.0 ? 2147483647 : t <= -2147483648.0 ? INT32_MIN : (int32_t)t;
                    }
                } else if (to == Dt::Bool) {
                    r.b = std::make_shared<std::vector<uint8_t>>(cnt);
                    for (size_t e = 0; e < cnt; e++) (*r.b)[e] = get(e) != 0;
                } else {
                    throw GraphError{};
                }
                break;
            }
            case Op::Concat: {
                if (n.in.empty()) throw GraphError{};
                const Val& first = in(v, n, 0);
                const int ax = norm_axis(n.axis, first.shape.size());
                std::vector<size_t> shape = first.shape;
                shape[ax] = 0;
                for (size_t k = 0; k < n.in.size(); k++) {
                    const Val& p = in(v, n, k);
                    if (p.dt != Dt::F32 || p.shape.size() != shape.size()) throw GraphError{};
                    for (size_t d = 0; d < shape.size(); d++)
                        if ((int)d != ax && p.shape[d] != first.shape[d]) throw GraphError{};
                    shape[ax] += p.shape[ax];
                }
                size_t outer = 1, inner = 1;
                for (int d = 0; d < ax; d++) outer *= shape[d];
                for (size_t d = ax + 1; d < shape.size(); d++) inner *= shape[d];
                std::vector<float> o;
                o.reserve(outer * shape[ax] * inner);
                for (size_t oi = 0; oi < outer; oi++)
                    for (size_t k = 0; k < n.in.size(); k++) {
                        const Val& p = in(v, n, k);
                        const size_t chunk = p.shape[ax] * inner;
                        for (size_t e = 0; e < chunk; e++) o.push_back(act_apply(n.act, (*p.f)[oi * chunk + e]));
                    }
                r = floats(shape, std::move(o));
                break;
            }
            case Op::FullyConnected: {
                const Val &x = in(v, n, 0), &w = in(v, n, 1);
                const Val* bias = (n.in.size() > 2 && n.in[2] >= 0) ? &in(v, n, 2) : nullptr;
                if (x.dt != Dt::F32 || w.dt != Dt::F32 || w.shape.size() != 2) throw GraphError{};
                const size_t nout = w.shape[0], nin = w.shape[1];
                if (nin == 0 || x.f->size() % nin) throw GraphError{};
                if (bias && (bias->dt != Dt::F32 || bias->f->size() != nout)) throw GraphError{};
                const size_t batch = x.f->size() / nin;
                std::vector<float> o(batch * nout);
                for (size_t rr = 0; rr < batch; rr++)
                    for (size_t oo = 0; oo < nout; oo++) {
                        const float* xr = x.f->data() + rr * nin;
                        const float* wr = w.f->data() + oo * nin;
                        float acc = dot_fma(xr, wr, nin);
                        if (bias) acc += (*bias->f)[oo];
                        o[rr * nout + oo] = act_apply(n.act, acc);
                    }
                std::vector<size_t> shape;
                if (n.keep_dims) {
                    shape = x.shape;
                    shape.back() = nout;
                } else {
                    shape = {batch, nout};
                }
                r = floats(shape, std::move(o));
                break;
            }
            case Op::BatchMatMul: {
                const Val &x = in(v, n, 0), &y = in(v, n, 1);
                if (x.dt != Dt::F32 || y.dt != Dt::F32 || x.shape.size() < 2 || y.shape.size() < 2) throw GraphError{};
                const size_t xr = x.shape.size(), yr = y.shape.size();
                const size_t m = n.adj_x ? x.shape[xr - 1] : x.shape[xr - 2];
                const size_t kk = n.adj_x ? x.shape[xr - 2] : x.shape[xr - 1];
                const size_t k2 = n.adj_y ? y.shape[yr - 1] : y.shape[yr - 2];
                const size_t nn = n.adj_y ? y.shape[yr - 2] : y.shape[yr - 1];
                if (kk != k2) throw GraphError{};
                std::vector<size_t> xb(x.shape.begin(), x.shape.end() - 2), yb(y.shape.begin(), y.shape.end() - 2);
                auto ob = bshape(xb, yb);
                std::vector<size_t> ix{0}, iy{0};  // no batch dimensions: one matrix each
                if (!ob.empty()) {
                    ix = bindex(xb, ob);
                    iy = bindex(yb, ob);
                }
                std::vector<float> o(ix.size() * m * nn);
                for (size_t bi = 0; bi < ix.size(); bi++) {
                    const float* xm = x.f->data() + ix[bi] * m * kk;
                    const float* ym = y.f->data() + iy[bi] * kk * nn;
                    if (kk == 1) {
                        // outer product: fma(p, q, +0) rounds p*q once and turns
                        // a -0 product into +0, which is exactly p*q + 0.0f
                        float* __restrict__ ob_ = o.data() + bi * m * nn;
                        const float* __restrict__ xa = xm;
                        if (nn == 2) {
                            const float y0 = ym[0], y1 = ym[1];
                            for (size_t a = 0; a < m; a++) {
                                ob_[2 * a] = xa[a] * y0 + 0.0f;
                                ob_[2 * a + 1] = xa[a] * y1 + 0.0f;
                            }
                        } else {
                            for (size_t a = 0; a < m; a++)
                                for (size_t c = 0; c < nn; c++) ob_[a * nn + c] = xa[a] * ym[c] + 0.0f;
                        }
                        continue;
                    }
                    for (size_t a = 0; a < m; a++)
                        for (size_t c = 0; c < nn; c++) {
                            float acc = 0.0f;
                            for (size_t l = 0; l < kk; l++) {
                                const float p = n.adj_x ? xm[l * m + a] : xm[a * kk + l];
                                const float q = n.adj_y ? ym[c * kk + l] : ym[l * nn + c];
                                acc = fmaf(p, q, acc);
                            }
                            o[bi * m * nn + a * nn + c] = acc;
                        }
                }
                if (ob.size() + 2 > 6) throw GraphError{};
                ob.push_back(m);
                ob.push_back(nn);
                r = floats(ob, std::move(o));
                break;
            }
            case Op::Reshape: {
                const Val& x = in(v, n, 0);
                std::vector<int32_t> spec;
                if (n.in.size() > 1 && n.in[1] >= 0) {
                    const Val& s = in(v, n, 1);
                    if (s.dt != Dt::I32) throw GraphError{};
                    spec = *s.i;
                } else if (n.has_shape) {
                    spec = n.new_shape;
                } else {
                    throw GraphError{};
                }
                if (spec.size() > 6) throw GraphError{};
                const size_t total = x.count();
                size_t known = 1;
                int infer = -1;
                std::vector<size_t> shape(spec.size());
                for (size_t k = 0; k < spec.size(); k++) {
                    if (spec[k] == -1) {
                        if (infer >= 0) throw GraphError{};
                        infer = (int)k;
                    } else if (spec[k] < 0) {
                        throw GraphError{};
                    } else {
                        shape[k] = (size_t)spec[k];
                        known *= shape[k];
                    }
                }
                if (infer >= 0) {
                    if (known == 0 || total % known) throw GraphError{};
                    shape[infer] = total / known;
                }
                r = x;
                r.shape = shape;
                if (r.count() != total) throw GraphError{};
                break;
            }
            case Op::ExpandDims: {
                const Val &x = in(v, n, 0), &a = in(v, n, 1);
                if (a.dt != Dt::I32 || a.i->size() != 1) throw GraphError{};
                const int rank = (int)x.shape.size();
                const int ax = (*a.i)[0] < 0 ? (*a.i)[0] + rank + 1 : (*a.i)[0];
                if (ax < 0 || ax > rank || rank >= 6) throw GraphError{};
                r = x;
                r.shape.insert(r.shape.begin() + ax, 1);
                break;
            }
            case Op::Shape: {
                const Val& x = in(v, n, 0);
                r.set = true;
                r.dt = Dt::I32;
                r.shape = {x.shape.size()};
                r.i = std::make_shared<std::vector<int32_t>>();
                for (size_t d : x.shape) r.i->push_back((int32_t)d);
                break;
            }
            case Op::Fill: {
                const Val &d = in(v, n, 0), &x = in(v, n, 1);
                if (d.dt != Dt::I32 || d.i->size() > 6 || x.count() != 1) throw GraphError{};
                r.set = true;
                r.dt = x.dt;
                for (int32_t k : *d.i) {
                    if (k < 0) throw GraphError{};
                    r.shape.push_back((size_t)k);
                }
                const size_t cnt = r.count();
                if (x.dt == Dt::F32) r.f = std::make_shared<std::vector<float>>(cnt, (*x.f)[0]);
                else if (x.dt == Dt::I32) r.i = std::make_shared<std::vector<int32_t>>(cnt, (*x.i)[0]);
                else r.b = std::make_shared<std::vector<uint8_t>>(cnt, (*x.b)[0]);
                break;
            }
            case Op::Split: {
                const Val &a = in(v, n, 0), &x = in(v, n, 1);
                if (a.dt != Dt::I32 || a.i->size() != 1 || x.dt != Dt::F32) throw GraphError{};
                const int ax = norm_axis((*a.i)[0], x.shape.size());
                const size_t parts = n.out.size();
                if (parts == 0 || x.shape[ax] % parts) throw GraphError{};
                std::vector<size_t> shape = x.shape;
                shape[ax] /= parts;
                size_t outer = 1, inner = 1;
                for (int d = 0; d < ax; d++) outer *= x.shape[d];
                for (size_t d = ax + 1; d < x.shape.size(); d++) inner *= x.shape[d];
                const size_t chunk = shape[ax] * inner;
                for (size_t p = 0; p < parts; p++) {
                    std::vector<float> o(outer * chunk);
                    for (size_t oi = 0; oi < outer; oi++)
                        memcpy(&o[oi * chunk], x.f->data() + (oi * parts + p) * chunk, sizeof(float) * chunk);
                    v[n.out[p]] = floats(shape, std::move(o));
                }
                return;
            }
            default: throw GraphError{};
        }
        v[n.out[0]] = std::move(r);
    }
};

void LatticeDeleter::operator()(LatticeModel* m) const { delete m; }

static Dt dt_of(int t) { return t == 0 ? Dt::F32 : t == 2 ? Dt::I32 : t == 6 ? Dt::Bool : Dt::None; }

// tflite.zig:232-360 and lattice.zig:23-41
std::unique_ptr<LatticeModel, LatticeDeleter> load_lattice(const std::string& bytes, int nb, std::string& err) {
    std::unique_ptr<LatticeModel, LatticeDeleter> m(new LatticeModel());
    try {
        Fb fb{(const uint8_t*)bytes.data(), bytes.size()};
        const size_t root = fb.at<uint32_t>(0);
        auto codes_v = fb.vec(root, 1), subgraphs = fb.vec(root, 2), buffers = fb.vec(root, 4);
        if (codes_v.second < 0 || subgraphs.second < 1 || buffers.second < 0) throw FbError{};
        const size_t g = fb.deref(subgraphs.first);
        std::vector<Op> codes(codes_v.second);
        for (long long i = 0; i < codes_v.second; i++) {
            const size_t t = fb.deref(codes_v.first + 4 * i);
            if (fb.field(t, 1)) { err = "UnsupportedOp"; return nullptr; }
            const int dep = fb.scalar<int8_t>(t, 0, 0), bi = fb.scalar<int32_t>(t, 3, 0);
            codes[i] = op_of(std::max(dep, bi));
        }
        auto tens = fb.vec(g, 0);
        if (tens.second < 0) throw FbError{};
        const size_t nt = (size_t)tens.second;
        std::vector<std::string> names(nt);
        m->dtypes.resize(nt);
        m->consts.resize(nt);
        for (size_t i = 0; i < nt; i++) {
            const size_t t = fb.deref(tens.first + 4 * i);
            names[i] = fb.str(t, 3);
            m->dtypes[i] = dt_of(fb.scalar<int8_t>(t, 1, 0));
            const uint32_t bidx = fb.scalar<uint32_t>(t, 2, 0);
            if (bidx == 0 || bidx >= (uint64_t)buffers.second) continue;
            const size_t bt = fb.deref(buffers.first + 4 * (size_t)bidx);
            const uint8_t* data = nullptr;
            size_t dlen = 0;
            auto dv = fb.vec(bt, 0);
            if (dv.second > 0) {
                if (dv.first > fb.n || fb.n - dv.first < (size_t)dv.second) throw FbError{};
                data = fb.p + dv.first;
                dlen = (size_t)dv.second;
            }
            if (dlen == 0) {
                const uint64_t off = fb.scalar<uint64_t>(bt, 1, 0), size = fb.scalar<uint64_t>(bt, 2, 0);
                if (off > 1 && size > 0) {
                    if (off > fb.n || fb.n - off < size) throw FbError{};
                    data = fb.p + off;
                    dlen = (size_t)size;
                }
            }
            if (dlen == 0) continue;
            Val c;
            c.set = true;
            c.dt = m->dtypes[i];
            auto sv = fb.vec(t, 0);
            if (sv.second > 6) { err = "UnsupportedType"; return nullptr; }
            for (long long d = 0; d < sv.second; d++) {
                const int32_t x = fb.at<int32_t>(sv.first + 4 * d);
                if (x < 0) throw FbError{};
                c.shape.push_back((size_t)x);
            }
            const size_t cnt = c.count();
            if (c.dt == Dt::None) { err = "UnsupportedType"; return nullptr; }
            if (c.dt == Dt::Bool) {
                if (dlen != cnt) throw FbError{};
                c.b = std::make_shared<std::vector<uint8_t>>(cnt);
                for (size_t e = 0; e < cnt; e++) (*c.b)[e] = data[e] != 0;
            } else {
                if (dlen != cnt * 4) throw FbError{};
                if (c.dt == Dt::F32) {
                    c.f = std::make_shared<std::vector<float>>(cnt);
                    memcpy(c.f->data(), data, 4 * cnt);
                } else {
                    c.i = std::make_shared<std::vector<int32_t>>(cnt);
                    memcpy(c.i->data(), data, 4 * cnt);
                }
            }
            m->consts[i] = c;
        }
        auto ops = fb.vec(g, 3);
        if (ops.second < 0) throw FbError{};
        auto tensor_list = [&](size_t t, int id) {
            std::vector<int> l;
            auto v = fb.vec(t, id);
            for (long long k = 0; k < v.second; k++) {
                const int32_t x = fb.at<int32_t>(v.first + 4 * k);
                if (x >= (int64_t)nt) throw FbError{};
                l.push_back(x);
            }
            return l;
        };
        for (long long i = 0; i < ops.second; i++) {
            const size_t t = fb.deref(ops.first + 4 * i);
            const uint32_t ci = fb.scalar<uint32_t>(t, 0, 0);
            if (ci >= codes.size()) throw FbError{};
            Node n;
            n.op = codes[ci];
            if (n.op == Op::Unknown) { err = "UnsupportedOp"; return nullptr; }
            n.in = tensor_list(t, 1);
            n.out = tensor_list(t, 2);
            for (int o : n.out)
                if (o < 0) throw FbError{};
            const size_t opt = fb.table(t, 4);
            auto act = [&](int id) {
                const int a = fb.scalar<int8_t>(opt, id, 0);
                if (a < 0 || a > 3) throw FbError{};
                return a;
            };
            if (opt) switch (n.op) {
                    case Op::Add: case Op::Sub: case Op::Mul: n.act = act(0); break;
                    case Op::Concat:
                        n.axis = fb.scalar<int32_t>(opt, 0, 0);
                        n.act = act(1);
                        break;
                    case Op::FullyConnected:
                        n.act = act(0);
                        if (fb.scalar<int8_t>(opt, 1, 0) != 0) { err = "UnsupportedOp"; return nullptr; }
                        n.keep_dims = fb.scalar<uint8_t>(opt, 2, 0) != 0;
                        break;
                    case Op::BatchMatMul:
                        n.adj_x = fb.scalar<uint8_t>(opt, 0, 0) != 0;
                        n.adj_y = fb.scalar<uint8_t>(opt, 1, 0) != 0;
                        break;
                    case Op::Reshape: {
                        auto v = fb.vec(opt, 0);
                        if (v.second >= 0) {
                            n.has_shape = true;
                            for (long long d = 0; d < v.second; d++) n.new_shape.push_back(fb.at<int32_t>(v.first + 4 * d));
                        }
                        break;
                    }
                    default: break;
                }
            m->nodes.push_back(std::move(n));
        }
        auto sigs = fb.vec(root, 7);
        for (long long s = 0; s < sigs.second; s++) {
            const size_t st = fb.deref(sigs.first + 4 * s);
            if (fb.str(st, 2) != "predict" && sigs.second > 1) continue;
            if (fb.scalar<uint32_t>(st, 4, 0) != 0) { err = "UnsupportedOp"; return nullptr; }
            for (int which = 0; which < 2; which++) {
                auto mv = fb.vec(st, which);
                for (long long k = 0; k < mv.second; k++) {
                    const size_t e = fb.deref(mv.first + 4 * k);
                    const uint32_t ti = fb.scalar<uint32_t>(e, 1, 0);
                    if (ti >= nt) throw FbError{};
                    (which == 0 ? m->inputs : m->outputs).emplace_back(fb.str(e, 0), (int)ti);
                }
            }
            break;
        }
        if (m->inputs.empty()) {
            for (int which = 1; which <= 2; which++) {
                auto v = fb.vec(g, which);
                for (long long k = 0; k < v.second; k++) {
                    const int32_t ti = fb.at<int32_t>(v.first + 4 * k);
                    if (ti < 0 || (size_t)ti >= nt) throw FbError{};
                    (which == 1 ? m->inputs : m->outputs).emplace_back(names[ti], ti);
                }
            }
        }
    } catch (const FbError&) {
        err = "InvalidModel";
        return nullptr;
    }
    static const char* kFeatures[4] = {"fvnsim", "fvnsim10_", "fstdnsim", "fvdegenergy"};
    m->nb = nb;
    for (int f = 0; f < 4; f++)
        for (int b = 0; b < nb; b++) {
            const int t = m->find(m->inputs, std::string(kFeatures[f]) + std::to_string(b));
            if (t < 0) { err = "MissingInput"; return nullptr; }
            m->feat.push_back(t);
        }
    m->tau = m->find(m->inputs, "tau");
    m->output = m->find(m->outputs, "predictions");
    if (m->tau < 0 || m->output < 0) { err = "MissingInput"; return nullptr; }
    m->plan_release();
    // One recording call on zero features fills the interpreter's broadcast
    // caches (shapes here do not depend on the feature values; each use
    // checks its shapes anyway) and gives the values the execution plan is
    // built from.
    {
        std::vector<Val> rec(m->consts.size());
        for (int t : m->feat) {
            rec[t].set = true;
            rec[t].dt = Dt::F32;
            rec[t].shape = {1};
            rec[t].f = std::make_shared<std::vector<float>>(1, 0.0f);
        }
        rec[m->tau].set = true;
        rec[m->tau].dt = Dt::F32;
        rec[m->tau].shape = {1};
        rec[m->tau].f = std::make_shared<std::vector<float>>(1, 0.5f);
        m->warming = true;
        bool ok = true;
        try {
            m->run(rec, true);
        } catch (const GraphError&) {
            ok = false;
        } catch (const FbError&) {
            ok = false;
        }
        m->warming = false;
        const char* e = getenv("VISQOL_HIP_LATTICE_PLAN");
        if (ok && !(e && atoi(e) == 0)) m->build_plan(rec);
    }
    return m;
}

bool lattice_predict(const LatticeModel& m, const double* f0, const double* f1, const double* f2, const double* f3,
                     double& out) {
    if (m.plan_ok) {
        thread_local std::vector<float> arena;
        if (arena.size() < m.plan_size) arena.resize(m.plan_size);
        float* A = arena.data();
        const double* fs[4] = {f0, f1, f2, f3};
        for (int f = 0; f < 4; f++)
            for (int b = 0; b < m.nb; b++) A[m.plan_feat[f * m.nb + b]] = (float)fs[f][b];
        m.run_plan(A);
        out = (double)A[m.plan_out];
        return true;
    }
    // Per-call values (constants are read in place) in a per-thread table
    // that every call leaves empty: released tensors are cleared as their
    // last reader finishes, the rest below.
    thread_local std::vector<Val> v;
    if (v.size() < m.consts.size()) v.resize(m.consts.size());
    const double* feats[4] = {f0, f1, f2, f3};
    for (int f = 0; f < 4; f++)
        for (int b = 0; b < m.nb; b++) {
            Val s;
            s.set = true;
            s.dt = Dt::F32;
            s.shape = {1};
            s.f = std::make_shared<std::vector<float>>(1, (float)feats[f][b]);
            v[m.feat[f * m.nb + b]] = s;
        }
    Val t;
    t.set = true;
    t.dt = Dt::F32;
    t.shape = {1};
    t.f = std::make_shared<std::vector<float>>(1, 0.5f);
    v[m.tau] = t;
    bool ok = true;
    try {
        m.run(v);
    } catch (const GraphError&) {
        ok = false;
    } catch (const FbError&) {
        ok = false;
    }
    if (ok) {
        const Val& o = v[m.output];
        ok = o.set && o.dt == Dt::F32 && !o.f->empty();
        if (ok) out = (double)(*o.f)[0];
        for (int t : m.unreleased) v[t] = Val{};
    } else {
        for (Val& x : v) x = Val{};
    }
    return ok;
}

}  // namespace vq
