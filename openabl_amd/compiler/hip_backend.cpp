// hip_backend.cpp -- the -b hip code generator.
//
// From an analysed script it writes, into the output directory:
//   main.cpp   the host program: the C backend's program (src/backend/
//              CPrinter.cpp:250-311 prints the same structs, RTTI tables,
//              globals, functions and main()), compiled with g++ so that
//              initialisation draws from the generator in gcc's argument
//              order, with simulate(...) replaced by the runtime's C ABI
//              (include/abl_hip.h): upload the agents, run the timesteps on
//              the GPU, download them in the reference's array order; save()
//              writes the same JSON text (abl_save_json);
//   steps.hip  the device program: every step function as a kernel over the
//              stepping agents (struct-of-arrays columns), its near loop over
//              the runtime's cell-sorted grid of the near agent type, in cell
//              order or, with hip.order=index, merged into the reference's
//              index order; run-time add()/removeCurrent(), random() inside
//              steps, the sequential step's reductions; the program
//              registers itself with the runtime when its library loads;
//   abl_gen.h  the host helpers main.cpp uses (asset/hip/abl_gen.h);
//   build.sh, run.sh.
// Expressions are printed by the C printer's rules (GenericCPrinter.cpp:
// 20-94): vector operators become helper calls with the C runtime's types,
// -v is v * -1.0, a non-integer % is fmod in double, folded constants print
// with 6 significant digits (GenericPrinter.cpp:39-52), so the generated
// program computes what the C backend's computes, operation by operation.
#include "hip_backend.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>

#include "fold.hpp"

namespace ablc {

namespace {

enum class Mode { Host, Device, Seq };

const std::set<std::string> &reserved() {
  static const std::set<std::string> r = {
      "alignas", "alignof", "and", "and_eq", "asm", "auto", "bitand", "bitor", "case", "catch",
      "char", "class", "compl", "const", "constexpr", "const_cast", "decltype", "default",
      "delete", "do", "double", "dynamic_cast", "enum", "explicit", "export", "extern", "float",
      "friend", "goto", "inline", "long", "mutable", "namespace", "noexcept", "not", "not_eq",
      "nullptr", "operator", "or", "or_eq", "private", "protected", "public", "register",
      "reinterpret_cast", "short", "signed", "sizeof", "static", "static_assert", "static_cast",
      "struct", "switch", "template", "this", "thread_local", "throw", "try", "typedef", "typeid",
      "typename", "union", "unsigned", "using", "virtual", "volatile", "wchar_t", "xor", "xor_eq",
      "F", "Prog", "Ctx", "c_", "a", "i", "abl", "gen", "vec2", "vec3", "agents", "agents_info",
      "sim_", "main", "abl_float", "float2", "float3", "printf", "exit", "fmod", "sin", "cos",
      "tan", "exp", "log", "pow", "sqrt", "round", "fmin", "fmax", "near_t"};
  return r;
}

std::string ident(const std::string &n) { return reserved().count(n) ? n + "_v" : n; }

// A float literal as the reference prints it: ostream's default 6
// significant digits, ".0" added when no '.' is shown (placed before an
// exponent, so that the text stays a C literal).
std::string float_lit(double v) {
  if (std::isnan(v)) return "__builtin_nan(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
  std::ostringstream s;
  s << v;
  std::string t = s.str();
  if (t.find('.') == std::string::npos) {
    const size_t e = t.find('e');
    if (e == std::string::npos) t += ".0";
    else t.insert(e, ".0");
  }
  return t;
}

std::string string_lit(const std::string &v) {
  std::string r = "\"";
  for (char c : v) {
    if (c == '"' || c == '\\') r += '\\';
    if (c == '\n') {
      r += "\\n";
      continue;
    }
    r += c;
  }
  return r + "\"";
}

struct Column {
  int member;
  int comp;  // 0..2 within a vector member
  std::string ctype;  // device element type ("F", "int32_t", "uint8_t")
};

// Statement/expression scan helpers ----------------------------------------------------

void each_expr(const Expr &e, const std::function<void(const Expr &)> &f) {
  f(e);
  for (auto &x : e.a) each_expr(*x, f);
  for (auto &m : e.inits) each_expr(*m.expr, f);
}

void each_stmt(const Stmt &s, const std::function<void(const Stmt &)> &fs,
               const std::function<void(const Expr &)> &fe) {
  fs(s);
  if (s.e0) each_expr(*s.e0, fe);
  if (s.e1) each_expr(*s.e1, fe);
  for (auto &b : s.body) each_stmt(*b, fs, fe);
  if (s.s0) each_stmt(*s.s0, fs, fe);
  if (s.s1) each_stmt(*s.s1, fs, fe);
}

class Gen {
 public:
  Gen(Script &s, const Analysis &a, const BackendContext &c) : S(s), A(a), ctx(c) {
    prog_name = c.input_name;
    auto it = c.config.find("hip.program");
    if (it != c.config.end()) prog_name = it->second;
    for (AgentDecl *ag : S.agents) {
      std::vector<Column> cols;
      for (size_t m = 0; m < ag->members.size(); m++) {
        const Type t = ag->members[m]->type.resolved;
        const int nc = t.is_vec() ? t.vec_len() : 1;
        std::string ct = t.is_bool() ? "uint8_t" : t.is_int() ? "int32_t" : "F";
        if (!(t.is_bool() || t.is_int() || t.is_float() || t.is_vec()))
          throw BackendError("agent member " + ag->name + "." + ag->members[m]->name +
                             " has a type the hip backend cannot store");
        for (int k = 0; k < nc; k++) cols.push_back({(int) m, k, ct});
      }
      if (cols.size() > 32)
        throw BackendError("agent " + ag->name + " has more than 32 member components");
      agent_cols[ag] = cols;
    }
  }

  std::string host_program();
  std::string device_program();

 private:
  Script &S;
  const Analysis &A;
  const BackendContext &ctx;
  std::string prog_name;
  std::map<const AgentDecl *, std::vector<Column>> agent_cols;

  // printing state
  Mode mode = Mode::Host;
  const FuncDecl *fn = nullptr;
  std::vector<char> loops;  // 'n' near loop (a lambda), 'l' other loop
  int ind = 0;
  int label = 0;
  std::ostringstream *out = nullptr;

  std::string nl() { return "\n" + std::string(2 * ind, ' '); }
  std::string fresh() { return "t" + std::to_string(++label) + "_"; }

  // ---- types ---------------------------------------------------------------------------
  std::string type_name(const Type &t) {
    switch (t.id) {
      case TI::Void: return "void";
      case TI::Bool: return "bool";
      case TI::Int: return "int";
      case TI::Float: return mode == Mode::Host ? "abl_float" : "F";
      case TI::String: return "const char *";
      case TI::Vec2: return mode == Mode::Host ? "float2" : "vec2<F>";
      case TI::Vec3: return mode == Mode::Host ? "float3" : "vec3<F>";
      case TI::Agent:
        if (!t.agent) throw BackendError("generic agent values are not supported");
        return mode == Mode::Host ? t.agent->name + " *" : t.agent->name + "<F>";
      default:
        throw BackendError("values of type " + [&] {
          std::ostringstream s;
          s << t;
          return s.str();
        }() + " are not supported by the hip backend");
    }
  }
  std::string vpre(const Type &t) {  // helper prefix of a vector type
    if (mode == Mode::Host) return t.id == TI::Vec2 ? "float2_" : "float3_";
    return t.id == TI::Vec2 ? "v2_" : "v3_";
  }
  std::string tmpl() { return mode == Mode::Host ? "" : "<F>"; }
  std::string ftype() { return mode == Mode::Host ? "abl_float" : "F"; }

  // ---- expressions -----------------------------------------------------------------------
  std::string expr(const Expr &e);
  std::string call(const Expr &e);
  std::string vec_op(BinOp op, const Expr &l, const Expr &r);

  // ---- statements ------------------------------------------------------------------------
  void stmt(const Stmt &s);
  void body(const std::vector<StmtP> &b) {
    for (auto &x : b) stmt(*x);
  }
  void block_or_stmt(const Stmt &s);
  void near_loop(const Stmt &s);
  void simulate(const Stmt &s);
  void add_in_step(const Expr &creation);

  // ---- pieces ----------------------------------------------------------------------------
  bool is_global_array(const Expr &e) {
    if (e.kind != EK::Var) return false;
    for (ConstDecl *c : S.consts)
      if (c->is_array && c->var.id == e.var.id) return true;
    return false;
  }
  const ConstDecl *global_of(const Expr &e) {
    if (e.kind != EK::Var) return nullptr;
    for (ConstDecl *c : S.consts)
      if (c->var.id == e.var.id) return c;
    return nullptr;
  }
  int col_index(const AgentDecl *ag, int member, int comp) {
    const auto &cs = agent_cols.at(ag);
    for (size_t k = 0; k < cs.size(); k++)
      if (cs[k].member == member && cs[k].comp == comp) return (int) k;
    return -1;
  }
  int member_index(const AgentDecl *ag, const std::string &name) {
    for (size_t m = 0; m < ag->members.size(); m++)
      if (ag->members[m]->name == name) return (int) m;
    return -1;
  }
  std::string load_member(const AgentDecl *ag, int m, const std::string &dst,
                          const std::string &cols, const std::string &slot);
  // a near-loop candidate's member from its record accessor nb_ (column k at
  // record position staged[k])
  std::string load_member_nb(const AgentDecl *ag, int m, const std::string &dst,
                             const std::vector<int> &staged);
  // the columns of the near agent a near loop reads: its position first
  std::vector<int> staged_cols(const AgentDecl *nag, const std::set<std::string> &need);
  // distances and lengths the device code only compares with constants:
  // kept as their float square sums, compared against exact bounds
  // (abl_root_bounds.hpp) instead of taking the root
  std::set<int> sq_vars;  // locals of the function being emitted
  // The innermost near loop whose test's square sum may stand for
  // dist(centre.pos, nx.pos) in its body: the loop variable, the centre
  // agent variable, the position member and the emitted SqLen's name.
  struct NearSq {
    int loop_var, centre_var;
    std::string pos, sq;
  };
  std::vector<NearSq> near_sq;
  // the SqLen of the enclosing near test when `call` is dist() of the loop
  // agent's and the centre's positions (either order), else ""
  std::string near_dist_sq(const Expr &call) const;
  static bool root_call(const Expr &e) {
    return e.kind == EK::Call && e.call == CallKind::Builtin &&
           (e.sig == "dist_float2" || e.sig == "dist_float3" || e.sig == "length_float2" ||
            e.sig == "length_float3");
  }
  bool const_expr(const Expr &e);
  void find_sq_vars(const FuncDecl &f);
  std::string root_compare(const Expr &e);
  // whether step f can run on the runtime's LDS tiles (its near loop as a
  // walk of each agent's stencil), and on its neighbour lists (opt-in)
  bool tile_capable(const FuncDecl &f);
  bool list_capable(const FuncDecl &f);
  std::string store_member(const AgentDecl *ag, int m, const std::string &src,
                           const std::string &cols, const std::string &slot);
  bool calls_rng(const FuncDecl &f) { return f.uses_rng; }
  std::string const_decl_host(const ConstDecl &c);
  std::string const_decl_device(const ConstDecl &c);
  std::string func_params(const FuncDecl &f);
  uint32_t step_writes(const FuncDecl &f, bool *copy_only);
  std::string agent_structs_device();
  std::string seq_step(const FuncDecl &f);
};

// ---- expressions ------------------------------------------------------------------------

std::string Gen::vec_op(BinOp op, const Expr &l, const Expr &r) {
  const Type v = l.type.is_vec() ? l.type : r.type;
  const std::string p = vpre(v);
  const bool host = mode == Mode::Host;
  std::string name;
  switch (op) {
    case BinOp::Add: name = host ? "add" : "add"; break;
    case BinOp::Sub: name = "sub"; break;
    case BinOp::Mul: name = host ? "mul_scalar" : "mul"; break;
    case BinOp::Div: name = host ? "div_scalar" : "div"; break;
    case BinOp::Eq: name = host ? "equals" : "eq"; break;
    case BinOp::Ne: name = host ? "not_equals" : "ne"; break;
    default: throw BackendError("vector operator not supported");
  }
  return p + name + tmpl() + "(" + expr(l) + ", " + expr(r) + ")";
}

std::string Gen::expr(const Expr &e) {
  switch (e.kind) {
    case EK::Var:
      return ident(e.var.name);
    case EK::IntLit:
      return std::to_string(e.ival);
    case EK::FloatLit:
      return float_lit(e.fval);
    case EK::BoolLit:
      return e.bval ? "true" : "false";
    case EK::StringLit:
      return string_lit(e.sval);
    case EK::Unary: {
      const Expr &x = *e.a[0];
      if (x.type.is_vec()) {
        if (e.uop == UnOp::Plus) return expr(x);
        return vpre(x.type) + (mode == Mode::Host ? "mul_scalar" : "mul") + tmpl() + "(" + expr(x) +
               ", -1.0)";
      }
      return std::string("(") + sigil(e.uop) + expr(x) + ")";
    }
    case EK::Binary: {
      const Expr &l = *e.a[0], &r = *e.a[1];
      if (l.type.is_vec() || r.type.is_vec()) return vec_op(e.bop, l, r);
      if (mode == Mode::Device) {
        const std::string rc = root_compare(e);
        if (!rc.empty()) return rc;
      }
      if (e.bop == BinOp::Mod && !(l.type.is_int() && r.type.is_int()))
        return "fmod((double) (" + expr(l) + "), (double) (" + expr(r) + "))";
      if (e.bop == BinOp::Range) throw BackendError("a range outside a for loop");
      return "(" + expr(l) + " " + sigil(e.bop) + " " + expr(r) + ")";
    }
    case EK::Ternary:
      return "(" + expr(*e.a[0]) + " ? " + expr(*e.a[1]) + " : " + expr(*e.a[2]) + ")";
    case EK::Member: {
      const Expr &x = *e.a[0];
      if (x.type.is_agent()) return expr(x) + (mode == Mode::Host ? "->" : ".") + e.name;
      if (x.type.is_vec()) return expr(x) + "." + e.name;
      throw BackendError("member access on an agent type outside a reduction");
    }
    case EK::Index: {
      const Expr &arr = *e.a[0];
      if (mode != Mode::Host && is_global_array(arr))
        return ident(arr.var.name) + "_at(" + expr(*e.a[1]) + ")";
      return expr(arr) + "[" + expr(*e.a[1]) + "]";
    }
    case EK::Call:
      return call(e);
    case EK::AgentNew: {
      // members initialised in the order written (a statement expression)
      AgentDecl *ag = e.type.agent;
      const std::string t = fresh();
      std::string s = "({ " + ag->name + tmpl() + " " + t + "; ";
      for (auto &mi : e.inits) s += t + "." + mi.name + " = " + expr(*mi.expr) + "; ";
      return s + t + "; })";
    }
    case EK::EnvAccess:
      throw BackendError("unresolved environment access");
    case EK::ArrayInit:
      throw BackendError("array initialisers are only supported in global constants");
    case EK::NewArray:
      throw BackendError("The hip backend does not support dynamic arrays (new T[n])");
  }
  return "";
}

std::string Gen::call(const Expr &e) {
  const bool host = mode == Mode::Host;
  auto args = [&](size_t from = 0) {
    std::string s;
    for (size_t k = from; k < e.a.size(); k++) s += (k > from ? ", " : "") + expr(*e.a[k]);
    return s;
  };
  if (e.call == CallKind::Ctor) {
    const Type t = e.type;
    if (t.is_vec()) {
      // arguments are converted to abl_float, as the C helpers' parameters do
      std::string s = vpre(t) + (e.a.size() == 1 ? (host ? "fill" : "fill") : "create") + tmpl() + "(";
      for (size_t k = 0; k < e.a.size(); k++)
        s += (k ? ", " : "") + std::string("(") + ftype() + ") (" + expr(*e.a[k]) + ")";
      return s + ")";
    }
    return "((" + type_name(t) + ") (" + expr(*e.a[0]) + "))";
  }
  if (e.call == CallKind::User) {
    std::string s = ident(e.sig) + "(";
    bool first = true;
    if (mode == Mode::Device && e.callee && e.callee->uses_rng) {
      s += "c_";
      first = false;
    }
    for (auto &x : e.a) {
      s += (first ? "" : ", ") + expr(*x);
      first = false;
    }
    return s + ")";
  }
  const std::string &n = e.sig;
  // vector geometry (asset/c/libabl.h:150-175)
  static const std::map<std::string, std::string> kGeom = {
      {"dot_float2", "dot2"},         {"dot_float3", "dot3"},         {"length_float2", "length2"},
      {"length_float3", "length3"},   {"dist_float2", "dist2"},       {"dist_float3", "dist3"},
      {"normalize_float2", "normalize2"}, {"normalize_float3", "normalize3"}};
  auto g = kGeom.find(n);
  if (!host && (n == "dist_float2" || n == "dist_float3")) {
    const std::string q = near_dist_sq(e);
    if (!q.empty()) return "(F) sqrtf_rn(" + q + ".s)";
  }
  if (g != kGeom.end()) return host ? n + "(" + args() + ")" : g->second + "<F>(" + args() + ")";
  if (n == "random_float") {
    if (mode == Mode::Device)
      return "c_.random_float((F) (" + expr(*e.a[0]) + "), (F) (" + expr(*e.a[1]) + "))";
    if (mode == Mode::Seq) throw BackendError("random() in a sequential step");
    return "random_float(" + args() + ")";
  }
  if (n == "random_int") {
    if (mode == Mode::Device) return "c_.random_int(" + args() + ")";
    if (mode == Mode::Seq) throw BackendError("randomInt() in a sequential step");
    return "random_int(" + args() + ")";
  }
  static const std::set<std::string> kMath = {"sin",  "cos",  "tan",  "sinh", "cosh",
                                               "tanh", "asin", "acos", "atan", "exp",
                                               "log",  "sqrt", "cbrt", "round"};
  if (kMath.count(n)) return "::" + n + "((double) (" + expr(*e.a[0]) + "))";
  if (n == "pow") return "::pow((double) (" + expr(*e.a[0]) + "), (double) (" + expr(*e.a[1]) + "))";
  if (n == "min" || n == "max") {
    if (host) return "abl_" + n + "(" + args() + ")";
    return (n == "min" ? "fmin_c<F>(" : "fmax_c<F>(") + std::string("(F) (") + expr(*e.a[0]) +
           "), (F) (" + expr(*e.a[1]) + "))";
  }
  if (n == "add") {
    if (mode == Mode::Host) {
      AgentDecl *ag = e.sig_params[0].agent;
      const Expr &arg = *e.a[0];
      return "(*(" + ag->name + " *) abl_place(&agents.agents_" + ag->name + ", sizeof(" +
             ag->name + ")) = " + (arg.kind == EK::AgentNew ? expr(arg) : "*" + expr(arg)) + ")";
    }
    throw BackendError("add() is only supported as a statement");
  }
  if (n == "save") {
    if (!host) throw BackendError("save() outside main()");
    return "abl_save_json(&agents, agents_info, " + expr(*e.a[0]) + ", (int) sizeof(abl_float))";
  }
  if (n == "removeCurrent") {
    if (mode != Mode::Device) throw BackendError("removeCurrent() outside a step");
    return "c_.remove_current()";
  }
  if (mode == Mode::Seq) {
    if (n == "count") return "gen::seq_count(sim_, " + std::to_string(e.a[0]->type.agent->index) + ")";
    if (n == "count_member" || n == "sum") {
      const Type mt = e.a[0]->type;
      const std::string q = std::to_string(mt.agent->index) + ", \"" + mt.member->name + "\"";
      if (n == "count_member")
        return "gen::seq_count_eq(sim_, " + q + ", (double) (" + expr(*e.a[1]) + "))";
      const Type r = mt.member->type.resolved;
      if (r.is_vec())
        return std::string(r.id == TI::Vec2 ? "gen::seq_sum2<F>" : "gen::seq_sum3<F>") + "(sim_, " +
               q + ")";
      return "((" + type_name(e.type) + ") gen::seq_sum(sim_, " + q + "))";
    }
    if (n == "getLastExecTime") return "((F) gen::seq_last_exec(sim_))";
    if (n == "log_csv") {
      std::string fmtstr, vals;
      for (auto &x : e.a) {
        fmtstr += (fmtstr.empty() ? "" : ",") + std::string(x->type.is_int() ? "%d" : "%f");
        vals += std::string(", ") + (x->type.is_int() ? "(int) (" : "(double) (") + expr(*x) + ")";
      }
      return "gen::seq_log(sim_, \"" + fmtstr + "\\n\"" + vals + ")";
    }
  }
  throw BackendError(e.name + "() is not supported here by the hip backend");
}

// ---- statements -------------------------------------------------------------------------

// The body of an if/while/for: braces on the same line.
void Gen::block_or_stmt(const Stmt &s) {
  *out << "{";
  ind++;
  if (s.kind == SK::Block) body(s.body);
  else stmt(s);
  ind--;
  *out << nl() << "}";
}

void Gen::stmt(const Stmt &s) {
  std::ostringstream &o = *out;
  switch (s.kind) {
    case SK::Expr: {
      const Expr &e = *s.e0;
      if (mode == Mode::Device && e.kind == EK::Call && e.sig == "add") {
        o << nl();
        add_in_step(*e.a[0]);
        return;
      }
      o << nl() << expr(e) << ";";
      return;
    }
    case SK::Assign: {
      const Expr &l = *s.e0, &r = *s.e1;
      if (mode == Mode::Host && r.type.is_agent()) o << nl() << "*" << expr(l) << " = *" << expr(r) << ";";
      else o << nl() << expr(l) << " = " << expr(r) << ";";
      return;
    }
    case SK::AssignOp: {
      const Expr &l = *s.e0, &r = *s.e1;
      if (l.type.is_vec() || r.type.is_vec()) {
        o << nl() << expr(l) << " = " << vec_op(s.op, l, r) << ";";
      } else if (s.op == BinOp::Mod && !(l.type.is_int() && r.type.is_int())) {
        o << nl() << expr(l) << " = fmod((double) (" << expr(l) << "), (double) (" << expr(r)
          << "));";
      } else {
        o << nl() << expr(l) << " " << sigil(s.op) << "= " << expr(r) << ";";
      }
      return;
    }
    case SK::Block:
      o << nl() << "{";
      ind++;
      body(s.body);
      ind--;
      o << nl() << "}";
      return;
    case SK::VarDecl: {
      const Type t = s.type.resolved;
      if (t.is_array()) throw BackendError("local arrays are not supported by the hip backend");
      if (mode == Mode::Host && t.is_agent()) {
        // storage plus a pointer, as the C printer declares agent variables
        const std::string st = fresh();
        o << nl() << t.agent->name << " " << st;
        if (s.e0) o << " = " << (s.e0->type.is_agent() ? "*" : "") << expr(*s.e0);
        o << ";" << nl() << t.agent->name << " *" << ident(s.var.name) << " = &" << st << ";";
        return;
      }
      if (mode == Mode::Device && sq_vars.count(s.var.id)) {
        // only compared with constants: the square sum, no root
        const Expr &c = *s.e0;
        std::string a;
        for (size_t k = 0; k < c.a.size(); k++) a += (k ? ", " : "") + expr(*c.a[k]);
        const std::string n = c.sig.substr(0, c.sig.find('_')) + (c.sig.back() == '2' ? "2" : "3");
        const std::string q = near_dist_sq(c);
        o << nl() << "const SqLen<F> " << ident(s.var.name) << " = "
          << (q.empty() ? "sq_" + n + "<F>(" + a + ")" : q) << ";";
        return;
      }
      o << nl() << type_name(t) << " " << ident(s.var.name);
      if (s.e0) o << " = " << expr(*s.e0);
      o << ";";
      return;
    }
    case SK::If:
      o << nl() << "if (" << expr(*s.e0) << ") ";
      block_or_stmt(*s.s0);
      if (s.s1) {
        o << " else ";
        block_or_stmt(*s.s1);
      }
      return;
    case SK::While:
      o << nl() << "while (" << expr(*s.e0) << ") ";
      loops.push_back('l');
      block_or_stmt(*s.s0);
      loops.pop_back();
      return;
    case SK::For: {
      if (s.for_kind == ForKind::Near) {
        near_loop(s);
        return;
      }
      if (s.for_kind != ForKind::Range)
        throw BackendError("for loops over arrays are not supported by the hip backend");
      const std::string e = fresh();
      const std::string v = ident(s.var.name);
      o << nl() << "for (int " << v << " = " << expr(*s.e0->a[0]) << ", " << e << " = "
        << expr(*s.e0->a[1]) << "; " << v << " < " << e << "; ++" << v << ") ";
      loops.push_back('l');
      block_or_stmt(*s.s0);
      loops.pop_back();
      return;
    }
    case SK::Simulate:
      simulate(s);
      return;
    case SK::Return: {
      const bool in_near = std::find(loops.begin(), loops.end(), 'n') != loops.end();
      if (in_near) {
        o << nl() << "{ ret_ = true; return false; }";
        return;
      }
      if (mode == Mode::Device && fn && fn->is_step()) {
        o << nl() << "return;";
        return;
      }
      if (s.e0) o << nl() << "return " << expr(*s.e0) << ";";
      else o << nl() << (mode == Mode::Seq && fn && fn->is_seq_step() ? "return 0;" : "return;");
      return;
    }
    case SK::Break:
      if (!loops.empty() && loops.back() == 'n') o << nl() << "return false;";
      else o << nl() << "break;";
      return;
    case SK::Continue:
      if (!loops.empty() && loops.back() == 'n') o << nl() << "return true;";
      else o << nl() << "continue;";
      return;
  }
}

std::string Gen::load_member(const AgentDecl *ag, int m, const std::string &dst,
                             const std::string &cols, const std::string &slot) {
  const auto &cs = agent_cols.at(ag);
  const Type t = ag->members[m]->type.resolved;
  std::string s;
  const char *comp[] = {"x", "y", "z"};
  for (size_t k = 0; k < cs.size(); k++) {
    if (cs[k].member != m) continue;
    const std::string target = dst + "." + ag->members[m]->name +
                               (t.is_vec() ? std::string(".") + comp[cs[k].comp] : "");
    if (t.is_bool())
      s += target + " = col_at<uint8_t>(" + cols + ", " + std::to_string(k) + ", " + slot + ") != 0; ";
    else
      s += target + " = col_at<" + cs[k].ctype + ">(" + cols + ", " + std::to_string(k) + ", " +
           slot + "); ";
  }
  return s;
}

std::string Gen::load_member_nb(const AgentDecl *ag, int m, const std::string &dst,
                                const std::vector<int> &staged) {
  const auto &cs = agent_cols.at(ag);
  const Type t = ag->members[m]->type.resolved;
  std::string s;
  const char *comp[] = {"x", "y", "z"};
  for (size_t k = 0; k < cs.size(); k++) {
    if (cs[k].member != m) continue;
    const std::string target = dst + "." + ag->members[m]->name +
                               (t.is_vec() ? std::string(".") + comp[cs[k].comp] : "");
    int slot = -1;
    for (size_t q = 0; q < staged.size(); q++)
      if (staged[q] == (int) k) slot = (int) q;
    const std::string args = "(" + std::to_string(k) + ", " + std::to_string(slot) + ")";
    if (t.is_bool())
      s += target + " = nb_.template get<uint8_t>" + args + " != 0; ";
    else
      s += target + " = nb_.template get<" + cs[k].ctype + ">" + args + "; ";
  }
  return s;
}

std::vector<int> Gen::staged_cols(const AgentDecl *nag, const std::set<std::string> &need) {
  const auto &cs = agent_cols.at(nag);
  const AgentMemberDecl *pos = nag->position();
  std::vector<int> r;
  for (size_t k = 0; k < cs.size(); k++)
    if (nag->members[cs[k].member].get() == pos) r.push_back((int) k);
  for (size_t k = 0; k < cs.size(); k++) {
    const AgentMemberDecl *md = nag->members[cs[k].member].get();
    if (md != pos && need.count(md->name)) r.push_back((int) k);
  }
  return r;
}

// A step runs on the runtime's LDS tiles when its one near loop walks the
// stepping agent's own type around the stepping agent itself, it neither
// adds nor removes agents, and the loop reads only float members of the
// candidates (the records staged in LDS are of the program's float type); on
// the neighbour lists too when the program is generated with -C hip.lists=1.
bool Gen::list_capable(const FuncDecl &f) {
  auto on = ctx.config.find("hip.lists");
  if (on == ctx.config.end() || (on->second != "1" && on->second != "true")) return false;
  return tile_capable(f);
}

bool Gen::tile_capable(const FuncDecl &f) {
  AgentDecl *ag = f.step_agent();
  if (!f.near_agent || f.near_agent != ag || f.uses_removal || f.added_agent) return false;
  int loops = 0;
  const Stmt *nl_stmt = nullptr;
  for (auto &st : f.body)
    each_stmt(
        *st,
        [&](const Stmt &x) {
          if (x.for_kind == ForKind::Near) {
            loops++;
            nl_stmt = &x;
          }
        },
        [](const Expr &) {});
  if (loops != 1) return false;
  const Expr &centre = *nl_stmt->e0->a[0];
  if (centre.kind != EK::Var || centre.var.id != f.params[0]->var.id) return false;
  std::set<std::string> need = f.near_members;
  need.insert(ag->position()->name);
  for (auto &md : ag->members) {
    if (!need.count(md->name)) continue;
    const Type t = md->type.resolved;
    if (!(t.is_float() || t.is_vec())) return false;
  }
  return true;
}

bool Gen::const_expr(const Expr &e) {
  bool constant = true;
  each_expr(e, [&](const Expr &x) {
    if (x.kind == EK::Var && !global_of(x)) constant = false;
    if (x.kind == EK::Member || (x.kind == EK::Call && x.call != CallKind::Ctor &&
                                 x.call != CallKind::Builtin))
      constant = false;
    if (x.kind == EK::Call && x.call == CallKind::Builtin &&
        (x.sig.rfind("random", 0) == 0 || x.sig == "count" || x.sig == "sum"))
      constant = false;
  });
  return constant && !e.type.is_vec();
}

std::string Gen::near_dist_sq(const Expr &call) const {
  if (call.kind != EK::Call || call.a.size() != 2 ||
      (call.sig != "dist_float2" && call.sig != "dist_float3"))
    return "";
  auto is_pos = [](const Expr &x, int var, const std::string &pos) {
    return x.kind == EK::Member && x.name == pos && !x.a.empty() && x.a[0]->kind == EK::Var &&
           x.a[0]->var.id == var;
  };
  const Expr &a = *call.a[0], &b = *call.a[1];
  for (auto it = near_sq.rbegin(); it != near_sq.rend(); ++it) {
    // (a - b) and (b - a) square to the same bits, summed in the same order
    if ((is_pos(a, it->loop_var, it->pos) && is_pos(b, it->centre_var, it->pos)) ||
        (is_pos(a, it->centre_var, it->pos) && is_pos(b, it->loop_var, it->pos)))
      return it->sq;
  }
  return "";
}

static bool relational(BinOp op) {
  return op == BinOp::Lt || op == BinOp::Le || op == BinOp::Gt || op == BinOp::Ge;
}

// A float local initialised with dist()/length(), never assigned again, and
// used only as an operand of <, <=, > or >= against a constant expression.
void Gen::find_sq_vars(const FuncDecl &f) {
  sq_vars.clear();
  std::map<int, int> uses, good;
  std::set<int> cand, assigned;
  std::function<void(const Expr &)> walk = [&](const Expr &x) {
    if (x.kind == EK::Var) uses[x.var.id]++;
    if (x.kind == EK::Binary && relational(x.bop)) {
      const Expr &l = *x.a[0], &r = *x.a[1];
      if (l.kind == EK::Var && const_expr(r)) good[l.var.id]++;
      else if (r.kind == EK::Var && const_expr(l)) good[r.var.id]++;
    }
    for (auto &y : x.a) walk(*y);
    for (auto &m : x.inits) walk(*m.expr);
  };
  for (auto &st : f.body)
    each_stmt(
        *st,
        [&](const Stmt &x) {
          if (x.kind == SK::VarDecl && x.e0 && x.type.resolved.is_float() && root_call(*x.e0))
            cand.insert(x.var.id);
          if ((x.kind == SK::Assign || x.kind == SK::AssignOp) && x.e0 && x.e0->kind == EK::Var)
            assigned.insert(x.e0->var.id);
          // the statement's own expressions (each_stmt's expression callback
          // would visit every subexpression, without the parent relation)
          if (x.e0 && !(x.kind == SK::Assign || x.kind == SK::AssignOp)) walk(*x.e0);
          if (x.kind == SK::Assign || x.kind == SK::AssignOp) {
            if (x.e0 && x.e0->kind != EK::Var) walk(*x.e0);
          }
          if (x.e1) walk(*x.e1);
        },
        [](const Expr &) {});
  for (int v : cand)
    if (!assigned.count(v) && uses[v] > 0 && uses[v] == good[v]) sq_vars.insert(v);
}

// `l op r` with a square-sum local or a dist()/length() call on one side and
// a constant on the other: the exact bound comparison (abl_root_bounds.hpp).
// The constant is converted as C does: an int to abl_float, a float operand
// compared in its own (double or promoted) value.
std::string Gen::root_compare(const Expr &e) {
  if (!relational(e.bop)) return "";
  const Expr &l = *e.a[0], &r = *e.a[1];
  auto sq_side = [&](const Expr &x) {
    return (x.kind == EK::Var && sq_vars.count(x.var.id)) || root_call(x);
  };
  const bool left = sq_side(l) && const_expr(r);
  const bool right = !left && sq_side(r) && const_expr(l);
  if (!left && !right) return "";
  const Expr &q = left ? l : r;
  const Expr &c = left ? r : l;
  BinOp op = e.bop;
  if (right) op = op == BinOp::Lt ? BinOp::Gt : op == BinOp::Gt ? BinOp::Lt
                : op == BinOp::Le ? BinOp::Ge : BinOp::Le;
  std::string qs;
  if (q.kind == EK::Var) {
    qs = ident(q.var.name);
  } else {
    std::string a;
    for (size_t k = 0; k < q.a.size(); k++) a += (k ? ", " : "") + expr(*q.a[k]);
    qs = near_dist_sq(q);
    if (qs.empty())
      qs = "sq_" + q.sig.substr(0, q.sig.find('_')) + (q.sig.back() == '2' ? "2" : "3") + "<F>(" + a + ")";
  }
  const std::string cv = c.type.is_int() ? "(double) (F) (" + expr(c) + ")" : "(double) (" + expr(c) + ")";
  const char *fnm = op == BinOp::Lt ? "len_lt" : op == BinOp::Le ? "len_le" : op == BinOp::Gt ? "len_gt" : "len_ge";
  return std::string(fnm) + "<F>(" + qs + ", " + cv + ")";
}

std::string Gen::store_member(const AgentDecl *ag, int m, const std::string &src,
                              const std::string &cols, const std::string &slot) {
  const auto &cs = agent_cols.at(ag);
  const Type t = ag->members[m]->type.resolved;
  std::string s;
  const char *comp[] = {"x", "y", "z"};
  for (size_t k = 0; k < cs.size(); k++) {
    if (cs[k].member != m) continue;
    const std::string from = src + "." + ag->members[m]->name +
                             (t.is_vec() ? std::string(".") + comp[cs[k].comp] : "");
    if (t.is_bool())
      s += "col_put<uint8_t>(" + cols + ", " + std::to_string(k) + ", " + slot + ", (uint8_t) (" +
           from + " ? 1 : 0)); ";
    else
      s += "col_put<" + cs[k].ctype + ">(" + cols + ", " + std::to_string(k) + ", " + slot + ", " +
           from + "); ";
  }
  return s;
}

// add(T { ... }) in a step: the newborn's columns at the parent's slot of
// the birth staging, and the parent's birth flag (the runtime appends the
// newborns after the step in ascending parent id).
void Gen::add_in_step(const Expr &creation) {
  AgentDecl *ag = creation.type.agent;
  const std::string t = fresh();
  *out << "{" << nl() << "  " << ag->name << "<F> " << t << " = " << expr(creation) << ";";
  *out << nl() << "  ";
  for (size_t m = 0; m < ag->members.size(); m++)
    *out << store_member(ag, (int) m, t, "c_.a.birth", "c_.i");
  *out << nl() << "  c_.a.born[c_.i] = 1;" << nl() << "}";
}

void Gen::near_loop(const Stmt &s) {
  if (mode != Mode::Device) throw BackendError("for-near loops are only supported in steps");
  std::ostringstream &o = *out;
  const Expr &call = *s.e0;
  AgentDecl *nag = s.type.resolved.agent;
  AgentMemberDecl *pos = nag->position();
  const bool dim3 = pos->type.resolved.id == TI::Vec3;
  // the centre: position of near()'s first argument
  const Expr &centre_agent = *call.a[0];
  const std::string p = fresh();
  bool has_return = false;
  each_stmt(
      *s.s0, [&](const Stmt &x) { has_return = has_return || x.kind == SK::Return; },
      [](const Expr &) {});
  o << nl() << "{";
  ind++;
  o << nl() << "const auto " << p << " = " << expr(centre_agent) << "." << pos->name << ";";
  if (has_return) o << nl() << "bool ret_ = false;";
  o << nl() << "c_.near(" << p << ".x, " << p << ".y, " << (dim3 ? p + ".z" : "(F) 0")
    << ", [&](const auto &nb_) -> bool {";
  ind++;
  const std::string v = ident(s.var.name);
  o << nl() << nag->name << "<F> " << v << ";";
  // the near agent's members the loop reads, and its position (from the
  // columns, or from the staged record of a list walk)
  std::set<std::string> need = fn ? fn->near_members : std::set<std::string>{};
  need.insert(pos->name);
  const std::vector<int> staged = staged_cols(nag, need);
  o << nl();
  for (size_t m = 0; m < nag->members.size(); m++)
    if (need.count(nag->members[m]->name)) o << load_member_nb(nag, (int) m, v, staged);
  // the near test on the square sum, kept for dist(nx.pos, centre.pos) in
  // the body while neither agent variable is assigned there
  const std::string nsq = fresh();
  o << nl() << "const SqLen<F> " << nsq << " = sq_dist" << (dim3 ? "3" : "2") << "<F>(" << v << "."
    << pos->name << ", " << p << ");";
  o << nl() << "if (c_.outside_sq(" << nsq << ")) return true;";
  bool reuse = centre_agent.kind == EK::Var;
  each_stmt(
      *s.s0,
      [&](const Stmt &x) {
        if ((x.kind != SK::Assign && x.kind != SK::AssignOp) || !x.e0) return;
        const Expr *t = x.e0.get();
        while ((t->kind == EK::Member || t->kind == EK::Index) && !t->a.empty()) t = t->a[0].get();
        if (t->kind == EK::Var && (t->var.id == s.var.id || t->var.id == centre_agent.var.id))
          reuse = false;
      },
      [](const Expr &) {});
  if (reuse) near_sq.push_back({s.var.id, centre_agent.var.id, pos->name, nsq});
  loops.push_back('n');
  if (s.s0->kind == SK::Block) body(s.s0->body);
  else stmt(*s.s0);
  loops.pop_back();
  if (reuse) near_sq.pop_back();
  o << nl() << "return true;";
  ind--;
  o << nl() << "});";
  if (has_return) o << nl() << "if (ret_) return;";
  ind--;
  o << nl() << "}";
}

void Gen::simulate(const Stmt &s) {
  if (mode != Mode::Host) throw BackendError("simulate outside main()");
  std::ostringstream &o = *out;
  o << nl() << "{";
  ind++;
  o << nl() << "abl_sim *sim_ = abl_gen_start(\"" << prog_name << "\", &agents, agents_info);";
  for (const auto &kv : ctx.config) {
    if (kv.first.rfind("hip.", 0) == 0 && kv.first != "hip.program" && kv.first != "hip.lists")
      o << nl() << "abl_gen_config(sim_, " << string_lit(kv.first) << ", " << string_lit(kv.second)
        << ");";
  }
  if (S.uses_logging) {
    // log_csv rows go to log.csv in the working directory, a new file per run
    o << nl() << "remove(\"log.csv\");";
    o << nl() << "abl_gen_config(sim_, \"hip.log_csv\", \"log.csv\");";
  }
  o << nl() << "abl_gen_config_env(sim_);";
  o << nl() << "abl_gen_run(sim_, (int) (" << expr(*s.e0) << "));";
  o << nl() << "abl_gen_finish(sim_, &agents, agents_info);";
  ind--;
  o << nl() << "}";
}

std::string Gen::func_params(const FuncDecl &f) {
  std::string s;
  for (size_t k = 0; k < f.params.size(); k++) {
    const Param &p = *f.params[k];
    s += (k ? ", " : "") + type_name(p.type.resolved) + " " + ident(p.var.name);
  }
  return s;
}

// Columns of the stepping agent a step writes (assignments to out.m or its
// components).  A member whose only assignment is out.m = in.m keeps its
// column (copy-through leaves it unchanged), so its grid stays valid.
uint32_t Gen::step_writes(const FuncDecl &f, bool *copy_only) {
  AgentDecl *ag = f.step_agent();
  const int out_id = f.params[0]->out->id, in_id = f.params[0]->var.id;
  std::map<int, int> assigns, copies;
  std::function<void(const Stmt &)> walk = [&](const Stmt &s) {
    if (s.kind == SK::Assign || s.kind == SK::AssignOp) {
      const Expr *root = s.e0.get();
      const Expr *top_member = nullptr;
      while (root->kind == EK::Member) {
        top_member = root;
        root = root->a[0].get();
      }
      if (root->kind == EK::Var && root->var.id == out_id) {
        if (!top_member) {
          for (size_t m = 0; m < ag->members.size(); m++) assigns[(int) m]++;
        } else {
          // the member of out: the Member node directly on the variable
          const Expr *m = s.e0.get();
          while (m->a[0]->kind == EK::Member) m = m->a[0].get();
          const int mi = member_index(ag, m->name);
          assigns[mi]++;
          const Expr &r = *s.e1;
          if (s.kind == SK::Assign && m == s.e0.get() && r.kind == EK::Member &&
              r.name == m->name && r.a[0]->kind == EK::Var && r.a[0]->var.id == in_id)
            copies[mi]++;
        }
      }
    }
    for (auto &b : s.body) walk(*b);
    if (s.s0) walk(*s.s0);
    if (s.s1) walk(*s.s1);
  };
  for (auto &s : f.body) walk(*s);
  uint32_t w = 0;
  *copy_only = false;
  const auto &cs = agent_cols.at(ag);
  for (auto &kv : assigns) {
    if (copies.count(kv.first) && copies[kv.first] == kv.second) {
      *copy_only = true;
      continue;
    }
    for (size_t k = 0; k < cs.size(); k++)
      if (cs[k].member == kv.first) w |= 1u << k;
  }
  return w;
}

std::string Gen::const_decl_host(const ConstDecl &c) {
  const Type t = c.type.resolved;
  std::ostringstream s;
  if (c.is_array) {
    const Type el = t.elem();
    s << type_name(el) << " " << ident(c.var.name) << "[] = { ";
    for (size_t k = 0; k < c.expr->a.size(); k++) s << (k ? ", " : "") << expr(*c.expr->a[k]);
    s << " };";
    return s.str();
  }
  if (t.is_vec()) {
    s << type_name(t) << " " << ident(c.var.name) << " = {";
    for (size_t k = 0; k < c.expr->a.size(); k++) s << (k ? ", " : "") << expr(*c.expr->a[k]);
    s << "};";
    return s.str();
  }
  s << type_name(t) << " " << ident(c.var.name) << " = " << expr(*c.expr) << ";";
  return s.str();
}

std::string Gen::const_decl_device(const ConstDecl &c) {
  const Type t = c.type.resolved;
  std::ostringstream s;
  if (c.is_array) {
    const Type el = t.elem();
    s << "static __host__ __device__ " << type_name(el) << " " << ident(c.var.name)
      << "_at(int k_) { const " << type_name(el) << " v_[] = { ";
    for (size_t k = 0; k < c.expr->a.size(); k++)
      s << (k ? ", " : "") << "(" << type_name(el) << ") " << expr(*c.expr->a[k]);
    s << " }; return v_[k_]; }";
    return s.str();
  }
  if (t.is_vec()) {
    s << "static constexpr " << type_name(t) << " " << ident(c.var.name) << " = {";
    for (size_t k = 0; k < c.expr->a.size(); k++) s << (k ? ", " : "") << "(F) " << expr(*c.expr->a[k]);
    s << "};";
    return s.str();
  }
  if (t.is_string()) {
    s << "static constexpr const char *" << ident(c.var.name) << " = " << expr(*c.expr) << ";";
    return s.str();
  }
  s << "static constexpr " << type_name(t) << " " << ident(c.var.name) << " = ("
    << type_name(t) << ") " << expr(*c.expr) << ";";
  return s.str();
}

// ---- the host program ----------------------------------------------------------------------

std::string Gen::host_program() {
  mode = Mode::Host;
  std::ostringstream o;
  out = &o;
  o << "// main.cpp -- generated by openabl-hip (-b hip) from " << ctx.input_name << ".abl.\n"
    << "// The C backend's program with simulate() run on the GPU through the runtime's C ABI.\n"
    << "#include \"abl_gen.h\"\n\n";
  // agent structs and their run-time type information
  static const char *kTypeId[] = {"", "", "ABL_TYPE_BOOL", "ABL_TYPE_INT", "ABL_TYPE_FLOAT",
                                  "ABL_TYPE_STRING", "ABL_TYPE_FLOAT2", "ABL_TYPE_FLOAT3"};
  for (AgentDecl *ag : S.agents) {
    o << "typedef struct {\n";
    for (auto &m : ag->members) o << "  " << type_name(m->type.resolved) << " " << m->name << ";\n";
    o << "} " << ag->name << ";\n";
    o << "static const abl_type_info " << ag->name << "_info[] = {\n";
    for (auto &m : ag->members)
      o << "  { " << kTypeId[(int) m->type.resolved.id] << ", offsetof(" << ag->name << ", "
        << m->name << "), \"" << m->name << "\", " << (m->is_position ? "true" : "false") << " },\n";
    o << "  { ABL_TYPE_END, sizeof(" << ag->name << "), NULL }\n};\n";
  }
  o << "struct agent_struct {\n";
  for (AgentDecl *ag : S.agents)
    o << "  abl_dyn_array agents_" << ag->name << ";\n  abl_dyn_array agents_" << ag->name
      << "_dbuf;\n";
  o << "};\nstruct agent_struct agents;\n";
  o << "static const abl_agent_info agents_info[] = {\n";
  for (AgentDecl *ag : S.agents)
    o << "  { " << ag->name << "_info, offsetof(struct agent_struct, agents_" << ag->name << "), \""
      << ag->name << "\" },\n";
  o << "  { NULL, 0, NULL }\n};\n\n";
  for (ConstDecl *c : S.consts) o << const_decl_host(*c) << "\n";
  o << "\n";
  // functions main() and its helpers use (not the step functions, which run on the GPU)
  for (FuncDecl *f : S.funcs) {
    if (f->is_any_step()) continue;
    fn = f;
    if (f->is_main()) {
      o << "int main() {";
      ind = 1;
      body(f->body);
      o << "\n  return 0;\n}\n";
      ind = 0;
      continue;
    }
    // a helper that only the GPU can evaluate is left out of the host program
    bool device_only = false;
    for (auto &s : f->body)
      each_stmt(
          *s, [&](const Stmt &x) { device_only = device_only || x.for_kind == ForKind::Near; },
          [&](const Expr &x) {
            device_only = device_only || (x.kind == EK::Call && x.sig == "removeCurrent");
          });
    if (device_only) continue;
    o << type_name(f->ret.resolved) << " " << ident(f->sig.name) << "(" << func_params(*f) << ") {";
    ind = 1;
    body(f->body);
    ind = 0;
    o << "\n}\n";
  }
  fn = nullptr;
  return o.str();
}

// ---- the device program ----------------------------------------------------------------------

std::string Gen::agent_structs_device() {
  std::ostringstream o;
  for (AgentDecl *ag : S.agents) {
    o << "template <typename F>\nstruct " << ag->name << " {\n";
    for (auto &m : ag->members) {
      const Type t = m->type.resolved;
      const std::string ty = t.is_bool() ? "bool" : t.is_int() ? "int" : t.is_float() ? "F"
                             : t.id == TI::Vec2 ? "vec2<F>" : "vec3<F>";
      o << "  " << ty << " " << m->name << ";\n";
    }
    o << "};\n";
  }
  return o.str();
}

std::string Gen::seq_step(const FuncDecl &f) {
  mode = Mode::Seq;
  fn = &f;
  std::ostringstream o;
  out = &o;
  o << "  // sequential step " << f.name << "() (once per timestep, on the host)\n";
  o << "  static int " << ident(f.name) << "(abl_sim *sim_) {";
  ind = 2;
  body(f.body);
  ind = 0;
  o << "\n    return 0;\n  }\n";
  mode = Mode::Device;
  return o.str();
}

std::string Gen::device_program() {
  mode = Mode::Device;
  std::ostringstream o;
  out = &o;
  const std::string ns = "gen_" + prog_name;
  o << "// steps.hip -- generated by openabl-hip (-b hip) from " << ctx.input_name << ".abl.\n"
    << "// Device program: one kernel per step function over the runtime's struct-of-arrays\n"
    << "// agents; registers program \"" << prog_name << "\" with the runtime when loaded.\n"
    << "#include <math.h>\n#include <stdio.h>\n\n#include \"abl_gen_device.hpp\"\n\n"
    << "namespace " << ns << " {\nusing namespace abl;\nusing namespace abl::gen;\n\n";
  o << agent_structs_device() << "\n";
  // per-agent column loads / stores
  for (AgentDecl *ag : S.agents) {
    o << "template <typename F>\n__device__ inline " << ag->name << "<F> load_" << ag->name
      << "(void *const *cols, int i) {\n  " << ag->name << "<F> r;\n  ";
    for (size_t m = 0; m < ag->members.size(); m++) o << load_member(ag, (int) m, "r", "cols", "i");
    o << "\n  return r;\n}\n";
  }
  o << "\ntemplate <typename F>\nstruct Prog {\n";
  for (ConstDecl *c : S.consts) o << "  " << const_decl_device(*c) << "\n";
  // functions: helpers (host and device unless they draw random numbers or
  // remove agents), step bodies
  for (FuncDecl *f : S.funcs) {
    if (f->is_main() || f->is_seq_step()) continue;
    fn = f;
    std::ostringstream fo;
    out = &fo;
    if (f->is_step()) {
      const Param &p = *f->params[0];
      fo << "  // step " << f->name << "\n  template <typename C>\n  static __device__ void "
         << ident(f->name) << "(C &c_, const " << p.type.resolved.agent->name << "<F> &"
         << ident(p.var.name) << ", " << p.type.resolved.agent->name << "<F> &"
         << ident(p.out->name) << ") {";
    } else if (f->uses_rng) {
      fo << "  template <typename C>\n  static __device__ " << type_name(f->ret.resolved) << " "
         << ident(f->sig.name) << "(C &c_" << (f->params.empty() ? "" : ", ") << func_params(*f)
         << ") {";
    } else {
      fo << "  static __host__ __device__ " << type_name(f->ret.resolved) << " " << ident(f->sig.name)
         << "(" << func_params(*f) << ") {";
    }
    ind = 2;
    find_sq_vars(*f);
    try {
      body(f->body);
    } catch (const BackendError &) {
      // a helper only main() uses may hold host-only calls (save, add):
      // leave it out of the device program; a step or a helper a step
      // calls fails the build instead
      if (f->is_step()) throw;
      ind = 0;
      loops.clear();
      continue;
    }
    ind = 0;
    fo << "\n  }\n";
    o << fo.str();
  }
  out = &o;
  fn = nullptr;
  // near radii (constant expressions) and the sequential step
  for (FuncDecl *f : S.simulate ? S.simulate->step_decls : std::vector<FuncDecl *>{}) {
    if (!f->near_agent) continue;
    const Stmt *nl_stmt = nullptr;
    for (auto &s : f->body)
      each_stmt(
          *s, [&](const Stmt &x) { if (x.for_kind == ForKind::Near) nl_stmt = &x; },
          [](const Expr &) {});
    const Expr *r = nl_stmt->e0->a[1].get();
    // a local initialised with a constant expression and never assigned
    // again stands for its initialiser, converted to its declared type
    std::string cast;
    if (r->kind == EK::Var && !global_of(*r)) {
      const Stmt *decl = nullptr;
      bool reassigned = false;
      for (auto &st : f->body)
        each_stmt(
            *st,
            [&](const Stmt &x) {
              if (x.kind == SK::VarDecl && x.var.id == r->var.id) decl = &x;
              if ((x.kind == SK::Assign || x.kind == SK::AssignOp) && x.e0 &&
                  x.e0->kind == EK::Var && x.e0->var.id == r->var.id)
                reassigned = true;
            },
            [](const Expr &) {});
      if (decl && decl->e0 && !reassigned) {
        cast = "(" + type_name(decl->type.resolved) + ") ";
        r = decl->e0.get();
      }
    }
    bool constant = true;
    each_expr(*r, [&](const Expr &x) {
      if (x.kind == EK::Var && !global_of(x)) constant = false;
      if (x.kind == EK::Member || (x.kind == EK::Call && x.call != CallKind::Ctor &&
                                   x.call != CallKind::Builtin))
        constant = false;
    });
    if (!constant)
      throw BackendError("The hip backend needs the near radius of " + f->name +
                         " to be a constant expression");
    o << "  static __host__ __device__ double radius_" << ident(f->name) << "() { return (double) ("
      << (cast.empty() ? expr(*r) : cast + "(" + expr(*r) + ")") << "); }\n";
  }
  if (S.simulate && S.simulate->seq_decl) o << seq_step(*S.simulate->seq_decl);
  o << "};\n\n";
  // kernels and launchers
  std::vector<FuncDecl *> steps = S.simulate ? S.simulate->step_decls : std::vector<FuncDecl *>{};
  for (FuncDecl *f : steps) {
    AgentDecl *ag = f->step_agent();
    bool copy_only = false;
    const uint32_t w = step_writes(*f, &copy_only);
    const std::string k = "k_" + f->name;
    o << "template <typename F, bool BY_ID>\n__global__ __launch_bounds__(256) void " << k
      << "(GenArgs a) {\n"
      << "  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);\n"
      << "  const int i = (int) (blk * blockDim.x + threadIdx.x);\n"
      << "  if (i >= a.n) return;\n"
      << "  Ctx<F, BY_ID> c_(a, i);\n"
      << "  const " << ag->name << "<F> in_ = load_" << ag->name << "<F>(a.in, i);\n"
      << "  " << ag->name << "<F> out_ = in_;\n"
      << "  Prog<F>::" << ident(f->name) << "(c_, in_, out_);\n  ";
    const auto &cs = agent_cols.at(ag);
    for (size_t m = 0; m < ag->members.size(); m++) {
      bool written = false;
      for (size_t c = 0; c < cs.size(); c++)
        if (cs[c].member == (int) m && ((w >> c) & 1u)) written = true;
      if (written) o << store_member(ag, (int) m, "out_", "a.out", "i");
    }
    o << "\n}\n";
    const bool tiles = tile_capable(*f);
    const bool lists = list_capable(*f);
    if (tiles) {
      // the step on the runtime's LDS tiles and lists (abl_gen_device.hpp)
      std::set<std::string> need = f->near_members;
      need.insert(ag->position()->name);
      const std::vector<int> staged = staged_cols(ag, need);
      AgentMemberDecl *pm = ag->position();
      const bool d3 = pm->type.resolved.id == TI::Vec3;
      o << "template <typename F_>\nstruct Step_" << f->name << " {\n"
        << "  using F = F_;\n  using Agent = " << ag->name << "<F>;\n"
        << "  static constexpr int DIM = " << (d3 ? 3 : 2) << ";\n"
        << "  static constexpr int NS = " << staged.size() << ";\n"
        << "  static constexpr int kCols[NS] = {";
      for (size_t q = 0; q < staged.size(); q++) o << (q ? ", " : "") << staged[q];
      o << "};\n"
        << "  static __device__ Agent load(const GenArgs &a, int i) { return load_" << ag->name
        << "<F>(a.in, i); }\n"
        << "  static __device__ void store(const GenArgs &a, int i, const Agent &out_) {\n    ";
      for (size_t m = 0; m < ag->members.size(); m++) {
        bool written = false;
        for (size_t c = 0; c < cs.size(); c++)
          if (cs[c].member == (int) m && ((w >> c) & 1u)) written = true;
        if (written) o << store_member(ag, (int) m, "out_", "a.out", "i");
      }
      o << "\n  }\n"
        << "  static __device__ void pos(const Agent &o, double *p) { p[0] = (double) o."
        << pm->name << ".x; p[1] = (double) o." << pm->name << ".y;"
        << (d3 ? " p[2] = (double) o." + pm->name + ".z;" : std::string()) << " }\n"
        << "  template <class C>\n  static __device__ void run(C &c_, const Agent &in, Agent &out) {\n"
        << "    Prog<F>::" << ident(f->name) << "(c_, in, out);\n  }\n};\n";
    }
    o << "template <typename F>\nvoid launch_" << f->name << "_t(const StepLaunch &L) {\n";
    if (lists)
      o << "  if (L.nl_mode) {\n    launch_gen_lists<Step_" << f->name << "<F>>(L);\n    return;\n  }\n";
    if (tiles)
      o << "  if (L.tile != 0 && L.order == 0 && L.grid.reach <= 4) {\n    launch_gen_tiles<Step_"
        << f->name << "<F>>(L);\n    return;\n  }\n";
    o << "  const GenArgs a = make_args(L);\n"
      << "  const unsigned nb = grid_blocks(L.n, 256);\n"
      << "  if (L.order == 1) hipLaunchKernelGGL((" << k << "<F, true>), nb, 256, 0, L.stream, a);\n"
      << "  else hipLaunchKernelGGL((" << k << "<F, false>), nb, 256, 0, L.stream, a);\n}\n"
      << "void launch_" << f->name << "(const StepLaunch &L) {\n"
      << "  if (L.precision == ABL_PREC_F32) launch_" << f->name << "_t<float>(L);\n"
      << "  else launch_" << f->name << "_t<double>(L);\n}\n";
    if (f->near_agent)
      o << "double radius_" << f->name << "(int precision) {\n"
        << "  return precision == ABL_PREC_F32 ? Prog<float>::radius_" << ident(f->name)
        << "() : Prog<double>::radius_" << ident(f->name) << "();\n}\n";
    o << "\n";
  }
  // environment (src/AnalysisVisitor.cpp:474-541, granularity :1378-1399)
  o << "void env(const double *, double lo[3], double hi[3], double *g) {\n";
  EnvDecl *env = S.env;
  if (env && env->dim > 0) {
    const auto mn = env->min.comps(), mx = env->max.comps();
    for (int d = 0; d < 3; d++)
      o << "  lo[" << d << "] = " << (d < env->dim ? float_lit(mn[d]) : "0.0") << "; hi[" << d
        << "] = " << (d < env->dim ? float_lit(mx[d]) : "0.0") << ";\n";
    std::ostringstream g;
    g.precision(17);
    g << env->granularity.as_float();
    o << "  *g = " << g.str() << ";\n";
  } else {
    o << "  for (int d = 0; d < 3; d++) lo[d] = hi[d] = 0.0;\n  *g = 1.0;\n";
  }
  o << "}\n";
  if (S.simulate && S.simulate->seq_decl) {
    o << "int seq_step(abl_sim *sim, int precision) {\n"
      << "  gen::seq_reset();\n"
      << "  const int rc = precision == ABL_PREC_F32 ? Prog<float>::"
      << ident(S.simulate->seq_decl->name) << "(sim) : Prog<double>::"
      << ident(S.simulate->seq_decl->name) << "(sim);\n"
      << "  return rc ? rc : gen::seq_status();\n}\n";
  }
  // the program description
  static const char *kKind[] = {"", "", "K_BOOL", "K_INT", "K_FLOAT", "", "K_FLOAT2", "K_FLOAT3"};
  o << "\nconst ProgramDesc *program() {\n  static const ProgramDesc d = {\n    \"" << prog_name
    << "\",\n    {";
  for (size_t t = 0; t < S.agents.size(); t++) {
    AgentDecl *ag = S.agents[t];
    o << (t ? ",\n     " : "") << "{\"" << ag->name << "\", {";
    for (size_t m = 0; m < ag->members.size(); m++) {
      auto &md = ag->members[m];
      o << (m ? ", " : "") << "{\"" << md->name << "\", " << kKind[(int) md->type.resolved.id]
        << ", " << (md->is_position ? "true" : "false") << "}";
    }
    o << "}}";
  }
  o << "},\n    {},\n    {";
  for (size_t k = 0; k < steps.size(); k++) {
    FuncDecl *f = steps[k];
    bool copy_only = false;
    const uint32_t w = step_writes(*f, &copy_only);
    const int adds = f->added_agent ? 1 + f->added_agent->index : 0;
    o << (k ? ",\n     " : "") << "{\"" << f->name << "\", " << f->step_agent()->index << ", "
      << (f->near_agent ? f->near_agent->index : -1) << ", nullptr, 0.0, " << w << "u, launch_"
      << f->name << ", " << (f->uses_removal ? "true" : "false") << ", " << adds << ", "
      << (list_capable(*f) ? "true" : "false") << ", "
      << (f->near_agent ? "radius_" + f->name : std::string("nullptr")) << "}";
  }
  o << "},\n    env,\n    nullptr,\n    nullptr,\n    nullptr,\n    false,\n    "
    << (S.simulate && S.simulate->seq_decl ? "seq_step" : "nullptr") << ",\n  };\n  return &d;\n}\n\n"
    << "const bool registered = register_program(program());\n\n}  // namespace " << ns << "\n";
  return o.str();
}

// Writes `text` unless the file already holds it (an unchanged file keeps
// its time stamp, so build.sh can tell that its outputs are up to date).
void write_file(const std::string &path, const std::string &text, bool exec = false) {
  {
    std::ifstream old(path);
    if (old) {
      std::ostringstream s;
      s << old.rdbuf();
      if (s.str() == text) {
        if (exec) ::chmod(path.c_str(), 0755);
        return;
      }
    }
  }
  std::ofstream f(path);
  if (!f) throw std::runtime_error("Could not open \"" + path + "\" for writing");
  f << text;
  f.close();
  if (exec) ::chmod(path.c_str(), 0755);
}

std::string read_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Could not read \"" + path + "\"");
  std::ostringstream s;
  s << f.rdbuf();
  return s.str();
}

}  // namespace

void HipBackend::generate(Script &script, const Analysis &analysis, const BackendContext &ctx) {
  if (!script.simulate) throw BackendError("The hip backend needs a simulate statement");
  for (FuncDecl *f : script.simulate->step_decls)
    if (!f->step_agent()) throw BackendError("step " + f->name + " has no agent type");
  Gen g(script, analysis, ctx);
  const std::string host = g.host_program();
  const std::string dev = g.device_program();
  const bool use_float = [&] {
    auto it = ctx.config.find("use_float");
    return it != ctx.config.end() && (it->second == "true" || it->second == "1");
  }();
  // the runtime: <root>/include, <root>/openabl_amd/csrc, <root>/openabl_amd/lib
  std::string root = ctx.asset_dir + "/../..";
  auto rit = ctx.config.find("hip.root");
  if (rit != ctx.config.end()) root = rit->second;
  const std::string d = ctx.output_dir;
  write_file(d + "/main.cpp", host);
  write_file(d + "/steps.hip", dev);
  write_file(d + "/abl_gen.h", read_file(ctx.asset_dir + "/hip/abl_gen.h"));
  std::ostringstream b;
  b << "#!/bin/sh\n# generated by openabl-hip: the device program, then the host program\n"
    << "set -e\nROOT=\"${ABL_HIP_ROOT:-" << root << "}\"\n"
    << "# nothing to do when both outputs are newer than their sources\n"
    << "if [ -f libsteps.so ] && [ -f main ] && [ libsteps.so -nt steps.hip ] && "
    << "[ main -nt main.cpp ] && [ main -nt abl_gen.h ]; then exit 0; fi\n"
    << "hipcc --offload-arch=${ABL_HIP_ARCH:-gfx950} -O3 -std=c++17 -ffp-contract=off "
    << "-Wno-parentheses-equality "
    << "-Xclang -target-feature -Xclang -packed-fp32-ops -fPIC -shared "
    << "-I\"$ROOT/include\" -I\"$ROOT/openabl_amd/csrc\" steps.hip "
    << "-L\"$ROOT/openabl_amd/lib\" -labl_hip -Wl,-rpath,\"$ROOT/openabl_amd/lib\" -o libsteps.so\n"
    << "# libsteps.so registers the program when it is loaded and main names none of\n"
    << "# its symbols: keep it linked where the linker defaults to --as-needed\n"
    << "g++ -O2 -std=gnu++17 -ffp-contract=off " << (use_float ? "-DABL_USE_FLOAT=1 " : "")
    << "-I\"$ROOT/include\" main.cpp -L. -Wl,--no-as-needed -lsteps -Wl,--as-needed "
    << "-L\"$ROOT/openabl_amd/lib\" -labl_hip "
    << "-Wl,-rpath,'$ORIGIN' -Wl,-rpath,\"$ROOT/openabl_amd/lib\" -lm -o main\n";
  write_file(d + "/build.sh", b.str(), true);
  write_file(d + "/run.sh", read_file(ctx.asset_dir + "/hip/run.sh"), true);
}

}  // namespace ablc
