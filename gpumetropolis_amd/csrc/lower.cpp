// lower.cpp -- reference bytecode -> HIP VM (see vm.h).
//
// The input is exactly what the reference's DSL compiler emits
// (R/dsl.R:28-122): (op, arg) pairs in postfix order with a deduplicated f64
// constant pool.  Evaluation semantics follow the reference interpreters
// (chain_kernel.rs:61-120, cpu_native.rs:80-107): binary ops pop r then l and
// push l op r, the result is the single value left on the stack, and an empty
// program evaluates to 0.
#include "lower.h"

#include <math.h>

#include <cmath>
#include <map>
#include <set>
#include <sstream>

#include "math64.h"

namespace gpum {

namespace {

VmInstr make_instr(uint32_t op, uint32_t slot = 0, uint32_t a = 0,
                   uint32_t b = 0) {
  VmInstr ins;
  ins.w0 = op | (slot << 8);
  ins.w1 = (a & 0xFFFFu) | (b << 16);
  return ins;
}

// Nominal VALU slots of one VM instruction per observation.
double vm_weight(uint32_t op) {
  switch (op) {
    case VM_LOAD_U: case VM_LOAD_D: case VM_PUSH_U: case VM_PUSH_D:
    case VM_SAVE: case VM_LOAD_S: case VM_PUSH_S:
      return 0.0;
    case VM_DIV_S: case VM_DIV_D: case VM_RDIV_U: case VM_RDIV_D:
    case VM_RCP:
      return 5.0;
    case VM_POW_S: case VM_POW_U: case VM_POW_D: case VM_RPOW_U:
    case VM_RPOW_D:
      return 9.0;
    case VM_EXP: case VM_LOG: case VM_SQRT: return 4.0;
    case VM_LGAMMA: return 32.0;
    case VM_CUBE: case VM_POW4: return 2.0;
    case VM_SOFTPLUS: return 9.0;
    // one EXP and one LOG, plus the max, difference and add
    case VM_LSE_S: return 11.0;
    default: return 1.0;
  }
}

// A uniform slot holding the constant v (one is added if none holds it).
int const_slot(Lowered* L, float v) {
  for (const auto& kv : L->const_uniforms)
    if (kv.second == v && std::signbit(kv.second) == std::signbit(v))
      return kv.first;
  const int slot = L->n_uniform++;
  L->const_uniforms.push_back({slot, v});
  return slot;
}

// Move work from the per-observation program into the summation (SumMode,
// vm.h).  Trailing ADD_U / SUB_U / MUL_U / FMA_UU go to the sum tail; a
// program
// left affine in the data columns, c + sum_i U[a_i] D[b_i] (heads LOAD_U c,
// LOAD_U a; MUL_D b, LOAD_D b; MUL_U a, or LOAD_D b; body FMA_UD, ADD_D,
// SUB_D, ADD_U, SUB_U),
// is rewritten as LOAD_U zero followed by one FMA_UD per product, summed
// by accumulation, with its constants at the front of the tail; otherwise a
// trailing SQR becomes the summation's square.
void fold_sum(Lowered* L) {
  std::vector<VmInstr>& P = L->vm;
  size_t end = P.size();
  while (end > 1) {
    const uint32_t op = vm_op(P[end - 1]);
    if (op != VM_ADD_U && op != VM_SUB_U && op != VM_MUL_U && op != VM_FMA_UU)
      break;
    --end;
  }
  std::vector<VmInstr> tail(P.begin() + (long)end, P.end());
  P.resize(end);

  // Affine form: products (a, b) and constants (ADD_U / SUB_U entries).
  std::vector<std::pair<int, int>> prods;
  std::vector<VmInstr> consts;
  bool affine = true;
  size_t pc = 1;
  const uint32_t op0 = vm_op(P[0]);
  if (op0 == VM_LOAD_U && P.size() > 1 && vm_op(P[1]) == VM_MUL_D) {
    prods.push_back({(int)vm_a(P[0]), (int)vm_a(P[1])});
    pc = 2;
  } else if (op0 == VM_LOAD_D && P.size() > 1 && vm_op(P[1]) == VM_MUL_U) {
    prods.push_back({(int)vm_a(P[1]), (int)vm_a(P[0])});
    pc = 2;
  } else if (op0 == VM_LOAD_U) {
    consts.push_back(make_instr(VM_ADD_U, 0, vm_a(P[0])));
  } else if (op0 == VM_LOAD_D) {
    prods.push_back({-1, (int)vm_a(P[0])});
  } else {
    affine = false;
  }
  for (; affine && pc < P.size(); ++pc) {
    const uint32_t op = vm_op(P[pc]);
    if (op == VM_FMA_UD)
      prods.push_back({(int)vm_a(P[pc]), (int)vm_b(P[pc])});
    else if (op == VM_ADD_D)
      prods.push_back({-1, (int)vm_a(P[pc])});
    else if (op == VM_SUB_D)
      prods.push_back({-2, (int)vm_a(P[pc])});
    else if (op == VM_ADD_U || op == VM_SUB_U)
      consts.push_back(make_instr(op, 0, vm_a(P[pc])));
    else
      affine = false;
  }
  if (affine) {
    std::vector<VmInstr> Q;
    Q.push_back(make_instr(VM_LOAD_U, 0, (uint32_t)const_slot(L, 0.0f)));
    for (auto& pr : prods) {
      int a = pr.first;
      if (a == -1) a = const_slot(L, 1.0f);
      if (a == -2) a = const_slot(L, -1.0f);
      Q.push_back(make_instr(VM_FMA_UD, 0, (uint32_t)a, (uint32_t)pr.second));
    }
    P.swap(Q);
    L->vm_depth = 1;
    L->sum_mode = SUM_AFFINE;
    tail.insert(tail.begin(), consts.begin(), consts.end());
  } else if (P.size() > 1 && vm_op(P.back()) == VM_SQR) {
    P.pop_back();
    L->sum_mode = SUM_SQUARE;
  }
  L->sum_tail = tail;
}

enum Kind { K_CONST = 0, K_PARAM = 1, K_DATA = 2, K_MIXED = 3 };

struct Node {
  uint32_t op = 0, arg = 0;
  int k0 = -1, k1 = -1;
  int kind = K_CONST;
  double val = 0.0;  // for K_CONST
  std::string key;   // structural key
};

bool is_unary(uint32_t op) {
  return op == OP_NEG || op == OP_EXP || op == OP_LOG || op == OP_SQRT ||
         op == OP_LGAMMA;
}
bool is_binary(uint32_t op) {
  return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
         op == OP_POW;
}

double apply_unary(uint32_t op, double x) {
  switch (op) {
    case OP_NEG: return -x;
    case OP_EXP: return exp(x);
    case OP_LOG: return log(x);
    case OP_SQRT: return sqrt(x);
    default: return lgamma64(x);
  }
}

double apply_binary(uint32_t op, double a, double b) {
  switch (op) {
    case OP_ADD: return a + b;
    case OP_SUB: return a - b;
    case OP_MUL: return a * b;
    case OP_DIV: return a / b;
    default: return pow(a, b);
  }
}

// Nominal VALU weights of the reference opcodes (unit = one FP32 lane op).
double w_unary(uint32_t op) {
  switch (op) {
    case OP_NEG: return 1.0;
    case OP_LGAMMA: return 32.0;
    default: return 4.0;  // EXP, LOG, SQRT
  }
}

struct Builder {
  std::vector<Node> nodes;
  int root = -1;
};

std::string build_tree(const uint32_t* code, size_t n_words,
                       const double* consts, Builder* b) {
  std::vector<int> st;
  for (size_t pc = 0; pc < n_words / 2; ++pc) {
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    Node n;
    n.op = op;
    n.arg = arg;
    std::ostringstream k;
    if (op == OP_PUSH_CONST) {
      n.kind = K_CONST;
      n.val = consts[arg];
      k.precision(17);
      k << "c" << n.val;
    } else if (op == OP_PUSH_PARAM) {
      n.kind = K_PARAM;
      k << "p" << arg;
    } else if (op == OP_PUSH_DATA) {
      n.kind = K_DATA;
      k << "d" << arg;
    } else if (is_unary(op)) {
      n.k0 = st.back();
      st.pop_back();
      const Node& c = b->nodes[n.k0];
      n.kind = c.kind;
      if (n.kind == K_CONST) n.val = apply_unary(op, c.val);
      k << "u" << op << "(" << c.key << ")";
    } else {
      n.k1 = st.back();
      st.pop_back();
      n.k0 = st.back();
      st.pop_back();
      const Node& l = b->nodes[n.k0];
      const Node& r = b->nodes[n.k1];
      int mask = (1 << l.kind) | (1 << r.kind);
      if (mask == (1 << K_CONST))
        n.kind = K_CONST;
      else if ((mask & ~((1 << K_CONST) | (1 << K_PARAM))) == 0)
        n.kind = K_PARAM;
      else if ((mask & ~((1 << K_CONST) | (1 << K_DATA))) == 0)
        n.kind = K_DATA;
      else
        n.kind = K_MIXED;
      if (n.kind == K_CONST) n.val = apply_binary(op, l.val, r.val);
      k << "b" << op << "(" << l.key << "," << r.key << ")";
    }
    n.key = k.str();
    b->nodes.push_back(n);
    st.push_back((int)b->nodes.size() - 1);
  }
  b->root = st.empty() ? -1 : st.back();
  return "";
}

// Weight of the program as compiled (nothing hoisted).  A MUL whose only
// consumer is an ADD or SUB is contracted into it (one FMA).
double bytecode_weight(const Builder& b) {
  std::vector<int> parent(b.nodes.size(), -1);
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    if (b.nodes[i].k0 >= 0) parent[b.nodes[i].k0] = (int)i;
    if (b.nodes[i].k1 >= 0) parent[b.nodes[i].k1] = (int)i;
  }
  double w = 0.0;
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const Node& n = b.nodes[i];
    if (n.op <= OP_PUSH_DATA) continue;
    if (is_unary(n.op)) {
      w += w_unary(n.op);
      continue;
    }
    const Node& r = b.nodes[n.k1];
    switch (n.op) {
      case OP_ADD:
      case OP_SUB:
        w += 1.0;
        break;
      case OP_MUL: {
        int p = parent[i];
        bool contracted =
            p >= 0 && (b.nodes[p].op == OP_ADD || b.nodes[p].op == OP_SUB);
        if (!contracted) w += 1.0;
        break;
      }
      case OP_DIV:
        w += (r.kind == K_CONST || r.kind == K_PARAM) ? 1.0 : 5.0;
        break;
      default: {  // POW
        if (r.kind == K_CONST && r.val == floor(r.val) && r.val >= 2.0 &&
            r.val <= 4.0)
          w += r.val == 2.0 ? 1.0 : 2.0;
        else
          w += 9.0;
      }
    }
  }
  return w;
}

enum Xf { XF_ID = 0, XF_NEG = 1, XF_RCP = 2, XF_NEGRCP = 3 };

struct Operand {
  bool is_u = true;
  int idx = 0;
};

struct Emitter {
  const Builder& b;
  const LowerOptions& opt;
  Lowered* out;
  std::map<std::string, int> uslot;
  std::map<std::string, int> dcol;
  int maxdepth = 0;
  std::string err;
  // Common subexpressions: repeated data-dependent subtrees evaluated once
  // and kept in a spill slot (numbered from 0 here, placed above the
  // program's stack by finish_cse).
  std::set<std::string> cse_keys;
  std::map<std::string, int> cse_slot;
  int n_cse = 0;

  Emitter(const Builder& bb, const LowerOptions& o, Lowered* L)
      : b(bb), opt(o), out(L) {}

  const Node& N(int i) const { return b.nodes[i]; }

  bool is_u(int i) const {
    return N(i).kind == K_CONST || N(i).kind == K_PARAM;
  }
  // Data-only subtrees are streamed as precomputed columns when their
  // per-observation cost is at least one transcendental (weight >= 4);
  // cheaper ones are recomputed by the VM rather than costing bandwidth.
  bool is_d(int i) const {
    const Node& n = N(i);
    if (n.kind != K_DATA) return false;
    if (n.op == OP_PUSH_DATA) return true;
    return opt.precompute_data && subtree_weight(i) >= 4.0;
  }
  double subtree_weight(int i) const {
    const Node& n = N(i);
    if (n.op <= OP_PUSH_DATA) return 0.0;
    double w = subtree_weight(n.k0);
    if (n.k1 >= 0) w += subtree_weight(n.k1);
    if (is_unary(n.op)) return w + w_unary(n.op);
    if (n.op == OP_DIV) return w + 5.0;
    if (n.op == OP_POW) return w + 9.0;
    return w + 1.0;
  }
  bool is_operand(int i) const { return is_u(i) || is_d(i); }
  bool is_const_val(int i, double v) const {
    return N(i).kind == K_CONST && N(i).val == v;
  }

  void hoist_postfix(int i) {
    const Node& n = N(i);
    if (n.k0 >= 0) hoist_postfix(n.k0);
    if (n.k1 >= 0) hoist_postfix(n.k1);
    out->hoist.push_back(n.op);
    out->hoist.push_back(n.arg);
  }

  void data_postfix(int i, std::vector<uint32_t>* code) {
    const Node& n = N(i);
    if (n.k0 >= 0) data_postfix(n.k0, code);
    if (n.k1 >= 0) data_postfix(n.k1, code);
    code->push_back(n.op);
    code->push_back(n.arg);
  }

  int uniform(int i, int xf) {
    std::string key = N(i).key + "#" + std::to_string(xf);
    auto it = uslot.find(key);
    if (it != uslot.end()) return it->second;
    int slot = out->n_uniform++;
    if (slot >= kMaxUniform) {
      err = "too many distinct hoisted values in the log-likelihood";
      return 0;
    }
    uslot[key] = slot;
    const Node& n = N(i);
    if (n.kind == K_CONST) {
      double v = n.val;
      if (xf & XF_RCP) v = 1.0 / v;
      if (xf & XF_NEG) v = -v;
      out->const_uniforms.push_back({slot, (float)v});
    } else {
      hoist_postfix(i);
      if (xf & XF_NEG) {
        out->hoist.push_back(OP_NEG);
        out->hoist.push_back(0);
      }
      out->hoist.push_back((xf & XF_RCP) ? OP_STORE_U_RCP : OP_STORE_U);
      out->hoist.push_back((uint32_t)slot);
    }
    return slot;
  }

  int column(int i) {
    const Node& n = N(i);
    if (n.op == OP_PUSH_DATA) return (int)n.arg;
    auto it = dcol.find(n.key);
    if (it != dcol.end()) return it->second;
    PrecomputedColumn pc;
    data_postfix(i, &pc.code);
    out->extra_cols.push_back(pc);
    int col = out->n_cols_in + (int)out->extra_cols.size() - 1;
    if (col >= kMaxCols) {
      err = "too many data columns after precomputation";
      return 0;
    }
    dcol[n.key] = col;
    return col;
  }

  Operand operand(int i, int xf = XF_ID) {
    Operand o;
    if (is_u(i)) {
      o.is_u = true;
      o.idx = uniform(i, xf);
    } else {
      o.is_u = false;
      o.idx = column(i);
    }
    return o;
  }

  void put(uint32_t op, uint32_t slot = 0, uint32_t a = 0, uint32_t bb = 0) {
    out->vm.push_back(make_instr(op, slot, a, bb));
  }

  void bump(int depth) {
    if (depth > maxdepth) maxdepth = depth;
  }

  // Repeated subtrees worth keeping: data-dependent, not an operand, at
  // least two VALU slots, marked at their outermost repetition (a repeat
  // inside one is evaluated once with it).
  void plan_cse(int root) {
    std::map<std::string, int> count;
    for (const Node& n : b.nodes) ++count[n.key];
    mark_cse(root, count);
  }
  void mark_cse(int i, const std::map<std::string, int>& count) {
    if (i < 0 || is_operand(i)) return;
    const Node& n = N(i);
    if (is_u(i)) return;
    if (count.at(n.key) >= 2 && subtree_weight(i) >= 2.0) {
      cse_keys.insert(n.key);
      return;
    }
    mark_cse(n.k0, count);
    mark_cse(n.k1, count);
  }
  // Place the CSE slots above the deepest spill slot the program uses.
  void finish_cse() {
    if (n_cse == 0) return;
    const uint32_t base = (uint32_t)std::max(0, maxdepth - 1);
    for (VmInstr& ins : out->vm) {
      const uint32_t op = vm_op(ins);
      if (op == VM_SAVE)
        ins = make_instr(op, base + vm_slot(ins), vm_a(ins), vm_b(ins));
      else if (op == VM_LOAD_S || op == VM_PUSH_S)
        ins = make_instr(op, vm_slot(ins), base + vm_a(ins), vm_b(ins));
    }
    maxdepth = (int)base + 1 + n_cse;
  }

  // Leave the value of node i in T; `d` values are on the stack below.
  void value(int i, int d) {
    const std::string& key = N(i).key;
    auto it = cse_slot.find(key);
    if (it != cse_slot.end()) {
      if (d == 0)
        put(VM_LOAD_S, 0, (uint32_t)it->second);
      else
        put(VM_PUSH_S, (uint32_t)(d - 1), (uint32_t)it->second);
      bump(d + 1);
      return;
    }
    value_raw(i, d);
    if (cse_keys.count(key)) {
      const int k = n_cse++;
      put(VM_SAVE, (uint32_t)k);
      cse_slot[key] = k;
    }
  }

  void value_raw(int i, int d) {
    if (is_operand(i)) {
      Operand o = operand(i);
      if (d == 0)
        put(o.is_u ? VM_LOAD_U : VM_LOAD_D, 0, o.idx);
      else
        put(o.is_u ? VM_PUSH_U : VM_PUSH_D, d - 1, o.idx);
      bump(d + 1);
      return;
    }
    emit(i, d);
  }

  // exp(A), c * exp(A) or exp(A) * c with a finite constant c > 0.
  bool match_cexp(int i, int* arg, double* c) const {
    const Node& n = N(i);
    if (n.op == OP_EXP) {
      *arg = n.k0;
      *c = 1.0;
      return true;
    }
    if (n.op != OP_MUL) return false;
    int e = -1, k = -1;
    if (N(n.k1).op == OP_EXP && N(n.k0).kind == K_CONST) {
      e = n.k1;
      k = n.k0;
    } else if (N(n.k0).op == OP_EXP && N(n.k1).kind == K_CONST) {
      e = n.k0;
      k = n.k1;
    }
    if (e < 0 || !(N(k).val > 0.0) || !std::isfinite(N(k).val)) return false;
    *arg = N(e).k0;
    *c = N(k).val;
    return true;
  }

  // MUL(U, D) or MUL(D, U): returns true and the operands.
  bool match_mul_ud(int i, int* u, int* dnode) const {
    const Node& n = N(i);
    if (n.op != OP_MUL) return false;
    if (is_u(n.k0) && is_d(n.k1)) { *u = n.k0; *dnode = n.k1; return true; }
    if (is_d(n.k0) && is_u(n.k1)) { *u = n.k1; *dnode = n.k0; return true; }
    return false;
  }

  // MUL(Z, U) or MUL(U, Z) with Z not a uniform.
  bool match_mul_zu(int i, int* z, int* u) const {
    const Node& n = N(i);
    if (n.op != OP_MUL) return false;
    if (is_u(n.k1) && !is_u(n.k0)) { *z = n.k0; *u = n.k1; return true; }
    if (is_u(n.k0) && !is_u(n.k1)) { *z = n.k1; *u = n.k0; return true; }
    return false;
  }

  void emit(int i, int d) {
    if (!err.empty()) return;
    const Node& n = N(i);
    bump(d + 1);
    if (is_unary(n.op)) {
      if (opt.fuse && n.op == OP_LOG) {
        const Node& a = N(n.k0);
        if (a.op == OP_ADD) {
          int e = -1;
          if (is_const_val(a.k0, 1.0) && N(a.k1).op == OP_EXP) e = a.k1;
          if (is_const_val(a.k1, 1.0) && N(a.k0).op == OP_EXP) e = a.k0;
          if (e >= 0) {
            value(N(e).k0, d);
            put(VM_SOFTPLUS);
            return;
          }
          // log(c1 exp(A1) + c2 exp(A2)), constants c > 0 (1 if absent):
          // log-sum-exp of A1 + log c1 and A2 + log c2, which cannot
          // underflow where both exponentials do in FP32 (mixture terms far
          // from a chain's means).  Equal weights add log c once, after.
          int a1 = -1, a2 = -1;
          double c1 = 1.0, c2 = 1.0;
          if (match_cexp(a.k0, &a1, &c1) && match_cexp(a.k1, &a2, &c2)) {
            const bool same = c1 == c2;
            value(a1, d);
            if (!same && c1 != 1.0)
              put(VM_ADD_U, 0, (uint32_t)const_slot(out, (float)std::log(c1)));
            value(a2, d + 1);
            if (!same && c2 != 1.0)
              put(VM_ADD_U, 0, (uint32_t)const_slot(out, (float)std::log(c2)));
            put(VM_LSE_S, (uint32_t)d);
            if (same && c1 != 1.0)
              put(VM_ADD_U, 0, (uint32_t)const_slot(out, (float)std::log(c1)));
            return;
          }
        }
      }
      value(n.k0, d);
      switch (n.op) {
        case OP_NEG: put(VM_NEG); break;
        case OP_EXP: put(VM_EXP); break;
        case OP_LOG: put(VM_LOG); break;
        case OP_SQRT: put(VM_SQRT); break;
        default: put(VM_LGAMMA); break;
      }
      return;
    }
    int L = n.k0, R = n.k1;
    // Small constant integer powers become multiplies.
    if (n.op == OP_POW && N(R).kind == K_CONST) {
      double e = N(R).val;
      if (e == 1.0 || e == 2.0 || e == 3.0 || e == 4.0 || e == -1.0 ||
          e == -2.0) {
        value(L, d);
        if (e == 2.0) put(VM_SQR);
        if (e == 3.0) put(VM_CUBE);
        if (e == 4.0) put(VM_POW4);
        if (e == -1.0) put(VM_RCP);
        if (e == -2.0) { put(VM_SQR); put(VM_RCP); }
        return;
      }
    }
    if (opt.fuse) {
      int u = -1, dn = -1, z = -1;
      // X +/- U*D  ->  FMA_UD
      if (n.op == OP_ADD || n.op == OP_SUB) {
        if (match_mul_ud(R, &u, &dn)) {
          int xf = n.op == OP_SUB ? XF_NEG : XF_ID;
          value(L, d);
          Operand ou = operand(u, xf), od = operand(dn);
          put(VM_FMA_UD, 0, ou.idx, od.idx);
          return;
        }
        if (n.op == OP_ADD && match_mul_ud(L, &u, &dn)) {
          value(R, d);
          Operand ou = operand(u), od = operand(dn);
          put(VM_FMA_UD, 0, ou.idx, od.idx);
          return;
        }
      }
      // Z*Ua +/- Ub, Ub +/- Z*Ua  ->  FMA_UU
      if (n.op == OP_ADD || n.op == OP_SUB) {
        int ua = -1;
        if (is_u(R) && match_mul_zu(L, &z, &ua)) {
          value(z, d);
          Operand oa = operand(ua);
          Operand ob = operand(R, n.op == OP_SUB ? XF_NEG : XF_ID);
          put(VM_FMA_UU, 0, oa.idx, ob.idx);
          return;
        }
        if (is_u(L) && match_mul_zu(R, &z, &ua)) {
          value(z, d);
          Operand oa = operand(ua, n.op == OP_SUB ? XF_NEG : XF_ID);
          Operand ob = operand(L);
          put(VM_FMA_UU, 0, oa.idx, ob.idx);
          return;
        }
      }
      // (-Z) * U, U * (-Z)  ->  MUL_U(-U)
      if (n.op == OP_MUL) {
        if (N(L).op == OP_NEG && !is_u(L) && is_u(R)) {
          value(N(L).k0, d);
          put(VM_MUL_U, 0, operand(R, XF_NEG).idx);
          return;
        }
        if (N(R).op == OP_NEG && !is_u(R) && is_u(L)) {
          value(N(R).k0, d);
          put(VM_MUL_U, 0, operand(L, XF_NEG).idx);
          return;
        }
      }
      // (-Z) / U  ->  MUL_U(-1/U)
      if (n.op == OP_DIV && N(L).op == OP_NEG && !is_u(L) && is_u(R)) {
        value(N(L).k0, d);
        put(VM_MUL_U, 0, operand(R, XF_NEGRCP).idx);
        return;
      }
    }
    // Z / U  ->  Z * (1/U), the reciprocal hoisted in FP64.
    if (n.op == OP_DIV && is_u(R) && !is_u(L)) {
      value(L, d);
      put(VM_MUL_U, 0, operand(R, XF_RCP).idx);
      return;
    }
    bool oL = is_operand(L), oR = is_operand(R);
    if (oR) {
      value(L, d);
      Operand o = operand(R);
      uint32_t op;
      switch (n.op) {
        case OP_ADD: op = o.is_u ? VM_ADD_U : VM_ADD_D; break;
        case OP_SUB: op = o.is_u ? VM_SUB_U : VM_SUB_D; break;
        case OP_MUL: op = o.is_u ? VM_MUL_U : VM_MUL_D; break;
        case OP_DIV:  // only D here (U handled above)
          if (o.is_u) {
            // L is a uniform too: T = U_l; T = T * (1/U_r)
            o = operand(R, XF_RCP);
            op = VM_MUL_U;
          } else {
            op = VM_DIV_D;
          }
          break;
        default: op = o.is_u ? VM_POW_U : VM_POW_D; break;
      }
      put(op, 0, o.idx);
      return;
    }
    if (oL) {
      value(R, d);
      Operand o = operand(L);
      uint32_t op;
      switch (n.op) {
        case OP_ADD: op = o.is_u ? VM_ADD_U : VM_ADD_D; break;
        case OP_SUB: op = o.is_u ? VM_RSUB_U : VM_RSUB_D; break;
        case OP_MUL: op = o.is_u ? VM_MUL_U : VM_MUL_D; break;
        case OP_DIV: op = o.is_u ? VM_RDIV_U : VM_RDIV_D; break;
        default: op = o.is_u ? VM_RPOW_U : VM_RPOW_D; break;
      }
      put(op, 0, o.idx);
      return;
    }
    value(L, d);
    value(R, d + 1);
    uint32_t op;
    switch (n.op) {
      case OP_ADD: op = VM_ADD_S; break;
      case OP_SUB: op = VM_SUB_S; break;
      case OP_MUL: op = VM_MUL_S; break;
      case OP_DIV: op = VM_DIV_S; break;
      default: op = VM_POW_S; break;
    }
    put(op, (uint32_t)d);
  }
};

}  // namespace

double eval_ref64(const uint32_t* code, size_t n_words, const double* consts,
                  const double* params, const double* row) {
  if (n_words < 2) return 0.0;
  double st[kMaxStack + 1];
  int sp = 0;
  for (size_t pc = 0; pc < n_words / 2; ++pc) {
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    switch (op) {
      case OP_PUSH_CONST: st[sp++] = consts[arg]; break;
      case OP_PUSH_PARAM: st[sp++] = params[arg]; break;
      case OP_PUSH_DATA: st[sp++] = row[arg]; break;
      case OP_ADD: st[sp - 2] += st[sp - 1]; --sp; break;
      case OP_SUB: st[sp - 2] -= st[sp - 1]; --sp; break;
      case OP_MUL: st[sp - 2] *= st[sp - 1]; --sp; break;
      case OP_DIV: st[sp - 2] /= st[sp - 1]; --sp; break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
      case OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case OP_LGAMMA: st[sp - 1] = lgamma64(st[sp - 1]); break;
      default: st[sp - 2] = pow(st[sp - 2], st[sp - 1]); --sp; break;
    }
  }
  return st[0];
}

std::string validate_program(const uint32_t* code, size_t n_words,
                             size_t n_consts, uint32_t n_params,
                             uint32_t n_cols, bool allow_data,
                             int* depth_out) {
  if (n_words % 2 != 0)
    return "bytecode length must be even (op, arg pairs)";
  if (n_words / 2 > 4096) return "bytecode program too long";
  int sp = 0, maxd = 0;
  for (size_t pc = 0; pc < n_words / 2; ++pc) {
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    if (op > OP_LGAMMA) {
      std::ostringstream m;
      m << "unknown opcode " << op << " at instruction " << pc;
      return m.str();
    }
    if (op == OP_PUSH_CONST) {
      if (arg >= n_consts) return "constant index out of range";
      ++sp;
    } else if (op == OP_PUSH_PARAM) {
      if (arg >= n_params) return "parameter index out of range";
      ++sp;
    } else if (op == OP_PUSH_DATA) {
      if (!allow_data) return "the prior must not read data columns";
      if (arg >= n_cols) return "data column index out of range";
      ++sp;
    } else if (is_unary(op)) {
      if (sp < 1) return "stack underflow in bytecode";
    } else {
      if (sp < 2) return "stack underflow in bytecode";
      --sp;
    }
    if (sp > maxd) maxd = sp;
    if (maxd > kMaxStack) {
      std::ostringstream m;
      m << "the log-density expression is too deep for the interpreter (needs "
        << maxd << " stack slots, limit " << kMaxStack << ")";
      return m.str();
    }
  }
  if (n_words > 0 && sp != 1)
    return "bytecode must leave exactly one value on the stack";
  if (depth_out) *depth_out = maxd;
  return "";
}

std::string lower_program(const uint32_t* code, size_t n_words,
                          const double* consts, size_t n_consts,
                          uint32_t n_params, uint32_t n_cols,
                          const LowerOptions& opt, Lowered* out) {
  *out = Lowered();
  out->n_cols_in = (int)n_cols;
  out->consts.assign(consts, consts + n_consts);
  std::string e = validate_program(code, n_words, n_consts, n_params, n_cols,
                                   true, &out->ref_depth);
  if (!e.empty()) return e;
  if (n_words == 0) {
    out->empty = true;
    return "";
  }
  Builder b;
  e = build_tree(code, n_words, consts, &b);
  if (!e.empty()) return e;
  out->w_bytecode = bytecode_weight(b);
  Emitter em(b, opt, out);
  if (opt.fuse) em.plan_cse(b.root);
  em.value(b.root, 0);
  if (!em.err.empty()) return em.err;
  em.finish_cse();
  out->hoist = square_pow2(out->hoist.data(), out->hoist.size(),
                           out->consts.data());
  out->vm_depth = em.maxdepth;
  if (out->vm_depth > kMaxStack) return "lowered program too deep";
  if (opt.fuse) fold_sum(out);
  out->w_exec = out->sum_mode == SUM_SQUARE ? 1.0 : 0.0;
  for (const VmInstr& ins : out->vm) out->w_exec += vm_weight(vm_op(ins));
  if ((int)(out->vm.size() + out->sum_tail.size()) > kMaxVmInstr)
    return "lowered program too long";
  return "";
}

std::vector<uint32_t> square_pow2(const uint32_t* code, size_t n_words,
                                  const double* consts) {
  std::vector<uint32_t> out;
  for (size_t pc = 0; pc + 1 < n_words; pc += 2) {
    const uint32_t op = code[pc], arg = code[pc + 1];
    if (op == OP_PUSH_CONST && pc + 3 < n_words && code[pc + 2] == OP_POW &&
        consts[arg] == 2.0) {
      out.push_back(OP_SQUARE);
      out.push_back(0);
      pc += 2;
      continue;
    }
    out.push_back(op);
    out.push_back(arg);
  }
  return out;
}

std::string disassemble(const Lowered& L) {
  static const char* names[] = {
      "LOAD_U", "LOAD_D", "PUSH_U", "PUSH_D", "ADD_S", "SUB_S", "MUL_S",
      "DIV_S", "POW_S", "ADD_U", "SUB_U", "MUL_U", "POW_U", "RSUB_U",
      "RDIV_U", "RPOW_U", "ADD_D", "SUB_D", "MUL_D", "DIV_D", "POW_D",
      "RSUB_D", "RDIV_D", "RPOW_D", "NEG", "EXP", "LOG", "SQRT", "LGAMMA",
      "SQR", "CUBE", "POW4", "RCP", "FMA_UU", "FMA_UD", "FMA_DU",
      "SOFTPLUS", "SAVE", "LOAD_S", "PUSH_S", "LSE_S"};
  std::ostringstream s;
  for (const VmInstr& i : L.vm) {
    uint32_t op = vm_op(i);
    s << (op < VM_NUM_OPS ? names[op] : "?") << " slot=" << vm_slot(i)
      << " a=" << vm_a(i) << " b=" << vm_b(i) << "\n";
  }
  static const char* modes[] = {"value", "square", "affine"};
  s << "sum: " << (L.sum_mode < 3 ? modes[L.sum_mode] : "?") << ", tail";
  for (const VmInstr& i : L.sum_tail) {
    uint32_t op = vm_op(i);
    s << " " << (op < VM_NUM_OPS ? names[op] : "?") << " " << vm_a(i);
  }
  s << "\nhoist:";
  for (size_t k = 0; k < L.hoist.size(); k += 2)
    s << " (" << L.hoist[k] << "," << L.hoist[k + 1] << ")";
  s << "\nn_uniform=" << L.n_uniform << " depth=" << L.vm_depth
    << " extra_cols=" << L.extra_cols.size() << " w_exec=" << L.w_exec
    << " w_bytecode=" << L.w_bytecode << "\n";
  return s.str();
}

}  // namespace gpum
