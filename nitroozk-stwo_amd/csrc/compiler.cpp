// Composition compiler: a component's constraint program (an expression DAG recorded through the
// constraint framework's EvalAtRow interface, nitroozk-stwo_amd/framework.py) becomes a HIP kernel
// at run time, so that evaluate_constraint_quotients_on_domain
// (crates/constraint-framework/src/prover/component_prover.rs:32-169) has a device kernel for any
// component, not only the hand-written ones (the reference dispatches by FNV-1a name and falls
// back to the CPU otherwise, component_prover.rs:409-496; its expression form is
// crates/constraint-framework/src/expr/).
//
// The generated kernel gives one lane one evaluation-domain row: it loads the row's mask values
// (the trace and interaction columns at the row and at offset rows of the trace domain), computes
// the DAG as straight-line code in M31 and QM31, accumulates sum_k pows[k] C_k (base-field
// constraints lazily in 64 bits), multiplies by the row's vanishing-polynomial inverse and adds the
// result into the accumulator, exactly as the hand-written kernels do. hipRTC compiles it for the
// current device once per distinct program; the module stays loaded for the process's lifetime.
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <map>
#include <mutex>
#include <sstream>


#include "compiler.h"

namespace sh {
namespace {

constexpr uint32_t MAGIC = 0x58454853u;
enum Op : uint32_t { OP_COL = 1, OP_CONST, OP_PARAM, OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_EXT };

// Field helpers of the generated source (plain integer code, no includes).
const char* kPrelude = R"(
typedef unsigned int u32;
typedef unsigned long long u64;
#define MP 0x7fffffffu
struct q4 { u32 a, b, c, d; };
struct acc4 { u64 a, b, c, d; };
__device__ __forceinline__ u32 madd(u32 a, u32 b) { u32 s = a + b; return s >= MP ? s - MP : s; }
__device__ __forceinline__ u32 msub(u32 a, u32 b) { u32 d = a - b, e = d + MP; return d < e ? d : e; }
__device__ __forceinline__ u32 mneg(u32 a) { return a ? MP - a : 0u; }
__device__ __forceinline__ u32 mred64(u64 x) {
  u64 r = (x & MP) + (x >> 31);
  u32 y = (u32)((r & MP) + (r >> 31));
  return y >= MP ? y - MP : y;
}
__device__ __forceinline__ u32 mmul(u32 a, u32 b) {
  u64 p = (u64)a * b;
  u32 s = (u32)(p & MP) + (u32)(p >> 31);
  return s >= MP ? s - MP : s;
}
__device__ __forceinline__ q4 qadd(q4 x, q4 y) { return {madd(x.a, y.a), madd(x.b, y.b), madd(x.c, y.c), madd(x.d, y.d)}; }
__device__ __forceinline__ q4 qsub(q4 x, q4 y) { return {msub(x.a, y.a), msub(x.b, y.b), msub(x.c, y.c), msub(x.d, y.d)}; }
__device__ __forceinline__ q4 qneg(q4 x) { return {mneg(x.a), mneg(x.b), mneg(x.c), mneg(x.d)}; }
__device__ __forceinline__ q4 qaddm(q4 x, u32 m) { return {madd(x.a, m), x.b, x.c, x.d}; }
__device__ __forceinline__ q4 qsubm(q4 x, u32 m) { return {msub(x.a, m), x.b, x.c, x.d}; }
__device__ __forceinline__ q4 msubq(u32 m, q4 x) { return {msub(m, x.a), mneg(x.b), mneg(x.c), mneg(x.d)}; }
__device__ __forceinline__ q4 qmulm(q4 x, u32 m) { return {mmul(x.a, m), mmul(x.b, m), mmul(x.c, m), mmul(x.d, m)}; }
__device__ __forceinline__ void cmul(u32 a, u32 b, u32 c, u32 d, u32& re, u32& im) {
  re = msub(mred64((u64)a * c), mred64((u64)b * d));
  im = mred64((u64)a * d + (u64)b * c);
}
__device__ __forceinline__ q4 qmul(q4 x, q4 y) {
  // (x0 + x1 u)(y0 + y1 u) = x0 y0 + (2 + i) x1 y1 + (x0 y1 + x1 y0) u
  u32 tr, ti, ar, ai, br, bi, cr, ci;
  cmul(x.c, x.d, y.c, y.d, tr, ti);
  cmul(x.a, x.b, y.a, y.b, ar, ai);
  cmul(x.a, x.b, y.c, y.d, br, bi);
  cmul(x.c, x.d, y.a, y.b, cr, ci);
  const u32 rr = msub(madd(tr, tr), ti), ri = madd(madd(ti, ti), tr);
  return {madd(ar, rr), madd(ai, ri), madd(br, cr), madd(bi, ci)};
}
__device__ __forceinline__ q4 qm(u32 m) { return {m, 0u, 0u, 0u}; }
__device__ __forceinline__ q4 qext(q4 e0, q4 e1, q4 e2, q4 e3) {
  return qadd(qadd(e0, qmul(e1, q4{0u, 1u, 0u, 0u})), qadd(qmul(e2, q4{0u, 0u, 1u, 0u}), qmul(e3, q4{0u, 0u, 0u, 1u})));
}
__device__ __forceinline__ void amad(acc4& s, q4 c, u32 v) {
  s.a += (u64)c.a * v; s.b += (u64)c.b * v; s.c += (u64)c.c * v; s.d += (u64)c.d * v;
}
__device__ __forceinline__ u64 f31(u64 p) { return (p & MP) + (p >> 31); }
__device__ __forceinline__ void afold(acc4& s) { s.a = f31(s.a); s.b = f31(s.b); s.c = f31(s.c); s.d = f31(s.d); }
__device__ __forceinline__ q4 ared(const acc4& s) { return {mred64(f31(s.a)), mred64(f31(s.b)), mred64(f31(s.c)), mred64(f31(s.d))}; }
__device__ __forceinline__ u32 brev(u32 i, u32 n) { return n ? (__builtin_bitreverse32(i) >> (32 - n)) : i; }
// offset_bit_reversed_circle_domain_index(row, trace_log, eval_log, off) (core/utils.rs): the row
// `off` trace steps away, in the evaluation domain's bit-reversed order
__device__ __forceinline__ u32 off_row(u32 row, u32 trace_log, u32 eval_log, int off) {
  const u32 p = brev(row, eval_log), half = 1u << (eval_log - 1);
  const int step = (int)(1u << (eval_log - trace_log - 1)) * off;
  u32 q;
  if (p < half) q = (u32)((int)p + step) & (half - 1);
  else q = ((u32)((int)(p - half) - step) & (half - 1)) + half;
  return brev(q, eval_log);
}
)";

struct Node { uint32_t op, a, b, c, d; bool secure; };
struct Program {
  std::vector<Node> nodes;
  std::vector<uint32_t> roots;
  uint32_t n_params = 0;
  uint32_t max_col[3] = {0, 0, 0};  // per tree: 1 + the largest column index read
  std::vector<int32_t> offsets;     // distinct non-zero row offsets
};

Program parse(const uint32_t* w, size_t n) {
  if (n < 5 || w[0] != MAGIC || w[1] != 1) throw HipError(1, "constraint program: bad header");
  Program p;
  const uint32_t n_nodes = w[2], n_cons = w[3];
  p.n_params = w[4];
  if (n_nodes > (1u << 20) || n_cons > (1u << 20) || p.n_params > 4096) throw HipError(3, "constraint program too large");
  if (n != 5 + 5 * (size_t)n_nodes + n_cons) throw HipError(1, "constraint program: length mismatch");
  for (uint32_t i = 0; i < n_nodes; ++i) {
    const uint32_t* x = w + 5 + 5 * (size_t)i;
    Node nd{x[0], x[1], x[2], x[3], x[4], false};
    auto arg = [&](uint32_t j) {
      if (j >= i) throw HipError(1, "constraint program: operand does not precede its node");
      return p.nodes[j].secure;
    };
    switch (nd.op) {
      case OP_COL:
        if (nd.a != 1 && nd.a != 2) throw HipError(1, "constraint program: column tree must be 1 or 2");
        p.max_col[nd.a] = std::max(p.max_col[nd.a], nd.b + 1);
        if ((int32_t)nd.c != 0 && std::find(p.offsets.begin(), p.offsets.end(), (int32_t)nd.c) == p.offsets.end())
          p.offsets.push_back((int32_t)nd.c);
        if ((int32_t)nd.c > (1 << 20) || (int32_t)nd.c < -(1 << 20)) throw HipError(1, "constraint program: offset");
        break;
      case OP_CONST:
        if (nd.a >= P) throw HipError(1, "constraint program: non-canonical constant");
        break;
      case OP_PARAM:
        if (nd.a > p.n_params) throw HipError(1, "constraint program: parameter index");
        nd.secure = true;
        break;
      case OP_ADD: case OP_SUB: case OP_MUL: {
        const bool sa = arg(nd.a), sb = arg(nd.b);  // both operands validated
        nd.secure = sa || sb;
        break;
      }
      case OP_NEG:
        nd.secure = arg(nd.a);
        break;
      case OP_EXT:
        arg(nd.a); arg(nd.b); arg(nd.c); arg(nd.d);
        nd.secure = true;
        break;
      default:
        throw HipError(1, "constraint program: unknown operation");
    }
    p.nodes.push_back(nd);
  }
  for (uint32_t k = 0; k < n_cons; ++k) {
    const uint32_t r = w[5 + 5 * (size_t)n_nodes + k];
    if (r >= n_nodes) throw HipError(1, "constraint program: constraint root");
    p.roots.push_back(r);
  }
  return p;
}

// Expression of node i's value from the names of its operands.
std::string node_expr(const Program& p, uint32_t i, const std::function<std::string(uint32_t)>& v,
                      const std::function<std::string(int32_t)>& rowname) {
  const Node& n = p.nodes[i];
  std::ostringstream s;
  const bool sa = n.op >= OP_ADD && n.op <= OP_NEG ? p.nodes[n.a].secure : false;
  const bool sb = n.op >= OP_ADD && n.op <= OP_MUL ? p.nodes[n.b].secure : false;
  auto asq = [&](uint32_t j) { return p.nodes[j].secure ? v(j) : "qm(" + v(j) + ")"; };
  switch (n.op) {
    case OP_COL: s << (n.a == 1 ? "trace[" : "inter[") << n.b << "][" << rowname((int32_t)n.c) << "]"; break;
    case OP_CONST: s << n.a << "u"; break;
    case OP_PARAM: s << "params[" << n.a << "]"; break;
    case OP_ADD:
      if (!sa && !sb) s << "madd(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa && sb) s << "qadd(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa) s << "qaddm(" << v(n.a) << ", " << v(n.b) << ")";
      else s << "qaddm(" << v(n.b) << ", " << v(n.a) << ")";
      break;
    case OP_SUB:
      if (!sa && !sb) s << "msub(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa && sb) s << "qsub(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa) s << "qsubm(" << v(n.a) << ", " << v(n.b) << ")";
      else s << "msubq(" << v(n.a) << ", " << v(n.b) << ")";
      break;
    case OP_MUL:
      if (!sa && !sb) s << "mmul(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa && sb) s << "qmul(" << v(n.a) << ", " << v(n.b) << ")";
      else if (sa) s << "qmulm(" << v(n.a) << ", " << v(n.b) << ")";
      else s << "qmulm(" << v(n.b) << ", " << v(n.a) << ")";
      break;
    case OP_NEG: s << (sa ? "qneg(" : "mneg(") << v(n.a) << ")"; break;
    case OP_EXT:
      if (!p.nodes[n.a].secure && !p.nodes[n.b].secure && !p.nodes[n.c].secure && !p.nodes[n.d].secure)
        s << "q4{" << v(n.a) << ", " << v(n.b) << ", " << v(n.c) << ", " << v(n.d) << "}";
      else
        s << "qext(" << asq(n.a) << ", " << asq(n.b) << ", " << asq(n.c) << ", " << asq(n.d) << ")";
      break;
  }
  return s.str();
}

// The kernel source. Large programs are cut into segments of at most SEG nodes, each a separate
// non-inlined function: the compiler's cost grows faster than linearly with the size of one
// straight-line function, and segments keep it linear in the program. Leaf nodes (columns,
// constants, parameters) are re-read in every segment that uses them; arithmetic values used in a
// later segment pass through per-lane slots, reused once their last use has passed. Each
// constraint is accumulated in the segment that computes it (base-field ones lazily in 64 bits,
// folded every 3 products and at every segment's end).
std::string gen_source(const Program& p) {
  constexpr uint32_t SEG = 1200;
  const uint32_t N = (uint32_t)p.nodes.size();
  auto leaf = [&](uint32_t i) { return p.nodes[i].op <= OP_PARAM; };
  std::vector<uint32_t> last_use(N, 0);
  auto use = [&](uint32_t j, uint32_t by) { last_use[j] = std::max(last_use[j], by); };
  for (uint32_t i = 0; i < N; ++i) {
    const Node& n = p.nodes[i];
    if (n.op >= OP_ADD && n.op <= OP_NEG) use(n.a, i);
    if (n.op >= OP_ADD && n.op <= OP_MUL) use(n.b, i);
    if (n.op == OP_EXT) { use(n.a, i); use(n.b, i); use(n.c, i); use(n.d, i); }
  }
  std::vector<std::vector<uint32_t>> roots_at(N);
  for (uint32_t k = 0; k < p.roots.size(); ++k) roots_at[p.roots[k]].push_back(k);
  const uint32_t n_seg = N ? (N + SEG - 1) / SEG : 1;
  // slots: per arithmetic value crossing a segment boundary, assigned when it is stored (end of its
  // segment) and freed after the segment of its last use
  std::vector<int> slot(N, -1);
  std::vector<int> free_m, free_q;
  uint32_t n_m = 0, n_q = 0;
  std::vector<std::vector<uint32_t>> stores(n_seg), frees(n_seg);
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t sd = i / SEG;
    if (!leaf(i) && last_use[i] > i && last_use[i] / SEG > sd) {
      stores[sd].push_back(i);
      frees[last_use[i] / SEG].push_back(i);
    }
  }
  std::ostringstream s;
  s << kPrelude;
  auto v = [](uint32_t i) { return "v" + std::to_string(i); };
  std::vector<std::string> segs;
  for (uint32_t g = 0; g < n_seg; ++g) {
    const uint32_t a = g * SEG, b = std::min(N, a + SEG);
    std::ostringstream f;
    f << "__device__ __attribute__((noinline)) void seg" << g
      << "(const u32* const* __restrict__ trace, const u32* const* __restrict__ inter,"
         " const q4* __restrict__ params, const q4* __restrict__ pows, const u32* rows, u32* lm, q4* lq,"
         " acc4& ab, q4& as) {\n  const u32 row = rows[0];\n";
    for (size_t o = 0; o < p.offsets.size(); ++o) f << "  const u32 r" << o << " = rows[" << o + 1 << "];\n";
    auto rowname = [&](int32_t off) -> std::string {
      if (off == 0) return "row";
      return "r" + std::to_string(std::find(p.offsets.begin(), p.offsets.end(), off) - p.offsets.begin());
    };
    // values from earlier segments: leaves re-read, arithmetic values from their slots
    std::vector<char> have(N, 0);
    std::function<void(uint32_t)> need = [&](uint32_t j) {
      if (j >= a || have[j]) return;
      have[j] = 1;
      const char* ty = p.nodes[j].secure ? "  const q4 " : "  const u32 ";
      if (leaf(j)) {
        f << ty << v(j) << " = " << node_expr(p, j, v, rowname) << ";\n";
      } else {
        if (slot[j] < 0) throw HipError(4, "constraint program: value without a slot");
        f << ty << v(j) << " = " << (p.nodes[j].secure ? "lq[" : "lm[") << slot[j] << "];\n";
      }
    };
    for (uint32_t i = a; i < b; ++i) {
      const Node& n = p.nodes[i];
      if (n.op >= OP_ADD && n.op <= OP_NEG) need(n.a);
      if (n.op >= OP_ADD && n.op <= OP_MUL) need(n.b);
      if (n.op == OP_EXT) { need(n.a); need(n.b); need(n.c); need(n.d); }
    }
    int pending = 0;
    for (uint32_t i = a; i < b; ++i) {
      f << (p.nodes[i].secure ? "  const q4 " : "  const u32 ") << v(i) << " = " << node_expr(p, i, v, rowname)
        << ";\n";
      for (uint32_t k : roots_at[i]) {
        if (p.nodes[i].secure) {
          f << "  as = qadd(as, qmul(pows[" << k << "], " << v(i) << "));\n";
        } else {
          f << "  amad(ab, pows[" << k << "], " << v(i) << ");\n";
          if (++pending == 3) { f << "  afold(ab);\n"; pending = 0; }
        }
      }
    }
    if (pending) f << "  afold(ab);\n";
    // slots: free the values whose last use is in this segment, then store this segment's outputs
    for (uint32_t j : frees[g]) (p.nodes[j].secure ? free_q : free_m).push_back(slot[j]);
    for (uint32_t j : stores[g]) {
      auto& fl = p.nodes[j].secure ? free_q : free_m;
      uint32_t& cnt = p.nodes[j].secure ? n_q : n_m;
      if (fl.empty()) slot[j] = (int)cnt++;
      else { slot[j] = fl.back(); fl.pop_back(); }
      f << (p.nodes[j].secure ? "  lq[" : "  lm[") << slot[j] << "] = " << v(j) << ";\n";
    }
    f << "}\n";
    segs.push_back(f.str());
  }
  for (auto& x : segs) s << x;
  s << "extern \"C\" __global__ void __launch_bounds__(256) expr_component(u32 eval_log, u32 trace_log,\n"
       "    const u32* const* __restrict__ trace, const u32* const* __restrict__ inter,\n"
       "    const q4* __restrict__ params, const q4* __restrict__ pows, const u32* __restrict__ dinv,\n"
       "    u32* acc0, u32* acc1, u32* acc2, u32* acc3) {\n"
       "  const u32 row = blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (row >= (1u << eval_log)) return;\n"
       "  u32 rows[" << p.offsets.size() + 1 << "];\n  rows[0] = row;\n";
  for (size_t o = 0; o < p.offsets.size(); ++o)
    s << "  rows[" << o + 1 << "] = off_row(row, trace_log, eval_log, " << p.offsets[o] << ");\n";
  s << "  u32 lm[" << std::max(1u, n_m) << "];\n  q4 lq[" << std::max(1u, n_q) << "];\n"
    << "  acc4 ab = {0ull, 0ull, 0ull, 0ull};\n  q4 as = {0u, 0u, 0u, 0u};\n";
  for (uint32_t g = 0; g < n_seg; ++g) s << "  seg" << g << "(trace, inter, params, pows, rows, lm, lq, ab, as);\n";
  s << "  q4 t = qmulm(qadd(ared(ab), as), dinv[row >> trace_log]);\n"
       "  t = qadd(q4{acc0[row], acc1[row], acc2[row], acc3[row]}, t);\n"
       "  acc0[row] = t.a; acc1[row] = t.b; acc2[row] = t.c; acc3[row] = t.d;\n}\n";
  return s.str();
}

struct Entry {
  std::vector<uint32_t> words;  // the program, to confirm a cache hit
  Program prog;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};
std::mutex g_mu;
std::multimap<std::pair<int, uint64_t>, std::unique_ptr<Entry>> g_cache;

uint64_t hash_words(const uint32_t* w, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= w[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

const Entry& get_compiled(int device, const uint32_t* w, size_t n) {
  const uint64_t h = hash_words(w, n);
  std::lock_guard<std::mutex> lk(g_mu);
  auto range = g_cache.equal_range({device, h});
  for (auto it = range.first; it != range.second; ++it)
    if (it->second->words.size() == n && std::memcmp(it->second->words.data(), w, n * 4) == 0) return *it->second;
  auto e = std::make_unique<Entry>();
  e->words.assign(w, w + n);
  e->prog = parse(w, n);
  const std::string src = gen_source(e->prog);
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, device));
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  if (auto c = arch.find(':'); c != std::string::npos) arch.resize(c);  // drop target features (xnack etc.)
  // code objects persist across processes in a cache directory, keyed by source and target
  std::string cache_path;
  {
    const std::string d = cache_directory();
    if (!d.empty()) {
      const std::string key = src + "\n" + arch;
      char name[40];
      std::snprintf(name, sizeof(name), "%016llx.co",
                    (unsigned long long)hash_words((const uint32_t*)key.data(), key.size() / 4) ^ key.size());
      cache_path = d + "/" + name;
      std::FILE* fp = std::fopen(cache_path.c_str(), "rb");
      if (fp) {
        std::vector<char> code;
        char buf[1 << 16];
        size_t k;
        while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) code.insert(code.end(), buf, buf + k);
        std::fclose(fp);
        if (!code.empty() && hipModuleLoadData(&e->mod, code.data()) == hipSuccess &&
            hipModuleGetFunction(&e->fn, e->mod, "expr_component") == hipSuccess) {
          const Entry& ref = *e;
          g_cache.emplace(std::make_pair(device, h), std::move(e));
          return ref;
        }
        e->mod = nullptr;
        e->fn = nullptr;
      }
    }
  }
  hiprtcProgram rp;
  if (hiprtcCreateProgram(&rp, src.c_str(), "expr_component.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw HipError(2, "hiprtcCreateProgram failed");
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(rp, 7, opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(rp, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(rp, &log[0]);
    hiprtcDestroyProgram(&rp);
    throw HipError(2, "constraint program compilation failed: " + log.substr(0, 400));
  }
  size_t cs = 0;
  hiprtcGetCodeSize(rp, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(rp, code.data());
  hiprtcDestroyProgram(&rp);
  if (!cache_path.empty()) {
    // written under a temporary name and renamed, so a concurrent reader never sees half a file
    const std::string tmp = cache_path + ".tmp" + std::to_string((unsigned long long)(uintptr_t)e.get());
    if (std::FILE* fp = std::fopen(tmp.c_str(), "wb")) {
      const bool ok = std::fwrite(code.data(), 1, code.size(), fp) == code.size();
      std::fclose(fp);
      if (!ok || std::rename(tmp.c_str(), cache_path.c_str()) != 0) std::remove(tmp.c_str());
    }
  }
  HIP_CHECK(hipModuleLoadData(&e->mod, code.data()));
  HIP_CHECK(hipModuleGetFunction(&e->fn, e->mod, "expr_component"));
  const Entry& ref = *e;
  g_cache.emplace(std::make_pair(device, h), std::move(e));
  return ref;
}

}  // namespace

uint32_t compiled_component_constraints(const uint32_t* words, size_t n_words) {
  if (n_words < 5 || words[0] != MAGIC) throw HipError(1, "constraint program: bad header");
  return words[3];
}

void evaluate_compiled_component(int device, const uint32_t* words, size_t n_words, uint32_t eval_log,
                                 uint32_t trace_log, const uint32_t* const* d_trace, uint32_t n_trace,
                                 const uint32_t* const* d_inter, uint32_t n_inter, const qm31* d_params,
                                 uint32_t n_params_given, const qm31* d_pows, uint32_t n_constraints,
                                 const uint32_t* d_dinv, Cols4 acc, hipStream_t stream) {
  const Entry& e = get_compiled(device, words, n_words);
  const Program& p = e.prog;
  if (p.roots.size() != n_constraints) throw HipError(1, "constraint program: constraint count differs");
  if (p.max_col[1] > n_trace || p.max_col[2] > n_inter) throw HipError(1, "constraint program reads a missing column");
  if (p.n_params + 1 > n_params_given) throw HipError(1, "constraint program: too few parameters");
  if (eval_log <= trace_log || eval_log > 30) throw HipError(3, "eval_log must exceed trace_log");
  const uint32_t* const* inter = d_inter ? d_inter : d_trace;
  uint32_t el = eval_log, tl = trace_log;
  uint32_t *a0 = acc.c[0], *a1 = acc.c[1], *a2 = acc.c[2], *a3 = acc.c[3];
  void* args[] = {&el, &tl, (void*)&d_trace, (void*)&inter, (void*)&d_params, (void*)&d_pows, (void*)&d_dinv,
                  &a0, &a1, &a2, &a3};
  const uint32_t n = 1u << eval_log;
  HIP_CHECK(hipModuleLaunchKernel(e.fn, (n + 255) / 256, 1, 1, 256, 1, 1, 0, stream, args, nullptr));
}

}  // namespace sh
