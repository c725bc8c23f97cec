// analysis.cpp -- semantic analysis of .abl scripts.
//
// One pass over the library script, then the main script, in declaration
// order, visiting each node's children before checking the node (a node's
// own checks run when it is left), as the reference's analysis does
// (src/AnalysisVisitor.cpp:22-1428).  Diagnostics carry the reference's
// texts and the order it reports them in, so that the lint corpus
// (test/*.abl, test/*.exp) reads the same.  What the analysis establishes for
// the code generator:
//   * types of every expression, with int -> float promotion as casts or
//     float literals, and scalar * vector rewritten as vector * scalar;
//   * folded global constants (their initialisers replaced by literals, -P
//     overrides substituted), environment bounds and granularity (the
//     largest constant near radius unless given);
//   * per function: its printed signature name (overloads get name_k, k the
//     number of earlier signatures of that name), its near loop's agent type
//     and the members the loop reads, runtime add()/removeCurrent(), random();
//   * the simulate statement's step and sequential-step functions.
// Two behaviours of the reference are kept because they decide results or
// diagnostics: a local declared with a global's name gets its own id but
// lookups of the name keep finding the global (the name map is only inserted
// into, never overwritten), and literals that replace folded expressions
// carry the default location (line 1).
#include "analysis.hpp"

#include <cassert>
#include <sstream>

#include "fold.hpp"

namespace ablc {

namespace {

Signature sig(const std::string &orig, const std::string &name, std::vector<Type> params, Type ret,
              unsigned flags = 0) {
  Signature s;
  s.orig = orig;
  s.name = name;
  s.params = std::move(params);
  s.ret = ret;
  s.flags = flags;
  return s;
}

std::string args_text(const std::vector<Type> &ts) {
  std::ostringstream s;
  s << "(";
  for (size_t k = 0; k < ts.size(); k++) s << (k ? ", " : "") << ts[k];
  s << ")";
  return s.str();
}

template <typename T>
std::string str(const T &t) {
  std::ostringstream s;
  s << t;
  return s.str();
}

bool compatible(const FunctionTable::Entry &e, const std::vector<Type> &args) {
  if (e.custom.compatible) return e.custom.compatible(args);
  if (args.size() != e.sig.params.size()) return false;
  for (size_t k = 0; k < args.size(); k++)
    if (!args[k].promotable_to(e.sig.params[k])) return false;
  return true;
}

bool conflicting(const Signature &s, const std::vector<Type> &params) {
  if (params.size() != s.params.size()) return false;  // arity overloading is allowed
  bool diff = false;
  for (size_t k = 0; k < params.size(); k++) {
    const Type &a = params[k], &b = s.params[k];
    if (a != b) {
      diff = true;
      const bool an = a.is_bool() || a.is_int() || a.is_float();
      const bool bn = b.is_bool() || b.is_int() || b.is_float();
      if (an && bn) return true;  // no overloading between bool, int and float
    }
  }
  return !diff;
}

Signature concrete(const FunctionTable::Entry &e, const std::vector<Type> &args) {
  if (e.custom.concrete) return e.custom.concrete(args);
  Signature s = e.sig;
  AgentDecl *agent = nullptr;
  for (size_t k = 0; k < s.params.size() && k < args.size(); k++) {
    if (s.params[k].is_agent() && !s.params[k].agent) {
      s.params[k] = args[k];
      if (!agent) agent = args[k].agent;
    }
  }
  if (s.ret.is_array() && s.ret.base == TI::Agent && !s.ret.agent) s.ret.agent = agent;
  return s;
}

AgentMemberDecl *find_member(AgentDecl &a, const std::string &name) {
  for (auto &m : a.members)
    if (m->name == name) return m.get();
  return nullptr;
}

bool ctor_valid(const Type &t, const std::vector<Type> &args) {
  const size_t n = args.size();
  auto all_float = [&] {
    for (const Type &a : args)
      if (!a.promotable_to(Type(TI::Float))) return false;
    return true;
  };
  switch (t.id) {
    case TI::Bool:
    case TI::Int:
    case TI::Float:
      return n == 1 && (args[0].is_bool() || args[0].is_num());
    case TI::Vec2:
      return (n == 1 || n == 2) && all_float();
    case TI::Vec3:
      return (n == 1 || n == 3) && all_float();
    default:
      return false;
  }
}

// `expr` is `<first parameter>.<pos>` of the current step function
bool is_in_pos(const Expr &e, const std::string &pos, const FuncDecl &f) {
  if (e.kind != EK::Member || e.name != pos) return false;
  const Expr &v = *e.a[0];
  return v.kind == EK::Var && !f.params.empty() && v.var.id == f.params[0]->var.id;
}

bool is_constant_expr(const Expr &e) {
  switch (e.kind) {
    case EK::IntLit:
    case EK::FloatLit:
    case EK::BoolLit:
    case EK::StringLit:
    case EK::Unary:
      return true;
    case EK::ArrayInit:
      for (const auto &x : e.a)
        if (!is_constant_expr(*x)) return false;
      return true;
    default:
      return false;
  }
}

}  // namespace

void register_builtins(FunctionTable &f) {
  const Type V(TI::Void), B(TI::Bool), I(TI::Int), F(TI::Float), V2(TI::Vec2), V3(TI::Vec3);
  const Type S(TI::String), A(TI::Agent), AT(TI::AgentType);
  f.add(sig("dot", "dot_float2", {V2, V2}, F));
  f.add(sig("dot", "dot_float3", {V3, V3}, F));
  f.add(sig("length", "length_float2", {V2}, F));
  f.add(sig("length", "length_float3", {V3}, F));
  f.add(sig("dist", "dist_float2", {V2, V2}, F));
  f.add(sig("dist", "dist_float3", {V3, V3}, F));
  f.add(sig("normalize", "normalize_float2", {V2}, V2));
  f.add(sig("normalize", "normalize_float3", {V3}, V3));
  f.add(sig("random", "random_float", {F, F}, F));
  f.add(sig("randomInt", "random_int", {I, I}, I));
  for (const char *m : {"sin", "cos", "tan", "sinh", "cosh", "tanh", "asin", "acos", "atan", "exp",
                        "log", "sqrt", "cbrt", "round"})
    f.add(sig(m, m, {F}, F));
  f.add(sig("pow", "pow", {F, F}, F));
  f.add(sig("min", "min", {F, F}, F));
  f.add(sig("max", "max", {F, F}, F));
  f.add(sig("add", "add", {A}, V, Signature::MainOnly | Signature::StepOnly));
  f.add(sig("removeCurrent", "removeCurrent", {}, V, Signature::StepOnly));
  Type near_ret = Type::array_of(A);
  f.add(sig("near", "near", {A, F}, near_ret, Signature::StepOnly));
  f.add(sig("save", "save", {S}, V, Signature::MainOnly));
  f.add(sig("count", "count", {AT}, I, Signature::SeqStepOnly));
  f.add(sig("getLastExecTime", "getLastExecTime", {}, F, Signature::SeqStepOnly));
  // sum(T.m): the member's type, bools summing to int
  Signature sum = sig("sum", "sum", {Type(TI::AgentMember)}, Type(TI::Unresolved),
                      Signature::SeqStepOnly);
  FunctionTable::Custom sc;
  sc.compatible = [](const std::vector<Type> &a) { return a.size() == 1 && a[0].is_agent_member(); };
  sc.concrete = [sum](const std::vector<Type> &a) {
    Signature s = sum;
    s.params = a;
    s.ret = a[0].member->type.resolved;
    if (s.ret.is_bool()) s.ret = Type(TI::Int);
    return s;
  };
  f.add(sum, sc);
  // count(T.m, v): how many agents have m == v
  Signature cm = sig("count", "count_member", {Type(TI::AgentMember), Type(TI::Unresolved)}, I,
                     Signature::SeqStepOnly);
  FunctionTable::Custom cc;
  cc.compatible = [](const std::vector<Type> &a) {
    if (a.size() != 2 || !a[0].is_agent_member()) return false;
    return a[1].compatible_with(a[0].member->type.resolved);
  };
  cc.concrete = [cm](const std::vector<Type> &a) {
    Signature s = cm;
    s.params = a;
    return s;
  };
  f.add(cm, cc);
  // log_csv(...): int and float arguments
  Signature lc = sig("log_csv", "log_csv", {}, V, Signature::SeqStepOnly);
  FunctionTable::Custom lcc;
  lcc.compatible = [](const std::vector<Type> &a) {
    for (const Type &t : a)
      if (!t.is_int() && !t.is_float()) return false;
    return true;
  };
  lcc.concrete = [lc](const std::vector<Type> &a) {
    Signature s = lc;
    s.params = a;
    return s;
  };
  f.add(lc, lcc);
}

Analysis::Analysis(Script &main, const std::map<std::string, std::string> &params,
                   Diagnostics &diag, const std::string &backend)
    : script_(main), params_(params), err_(diag), backend_(backend) {
  register_builtins(funcs_);
}

void Analysis::library(Script &lib) {
  lib_ = true;
  run(lib);
}

void Analysis::main_script() {
  lib_ = false;
  run(script_);
  finish();
}

void Analysis::run(Script &s) {
  for (Decl &d : s.decls) {
    switch (d.kind) {
      case Decl::Agent: agent(*d.agent); break;
      case Decl::Func:
        d.func->from_lib = lib_;
        func(*d.func);
        break;
      case Decl::Const:
        d.cnst->from_lib = lib_;
        cnst(*d.cnst);
        break;
      case Decl::Env: env(*d.env); break;
    }
  }
}

Type Analysis::try_simple(const std::string &n) const {
  if (n == "void") return Type(TI::Void);
  if (n == "bool") return Type(TI::Bool);
  if (n == "int") return Type(TI::Int);
  if (n == "float") return Type(TI::Float);
  if (n == "string") return Type(TI::String);
  if (n == "float2") return Type(TI::Vec2);
  if (n == "float3") return Type(TI::Vec3);
  return Type();
}

Type Analysis::resolve(TypeRef &t) {
  Type r = try_simple(t.name);
  if (r.invalid()) {
    auto it = agents_.find(t.name);
    if (it == agents_.end()) {
      err_.error("Unknown type \"" + t.name + "\"", t.loc);
      return Type();
    }
    r = Type::agent_of(TI::Agent, it->second);
  }
  return r;
}

void Analysis::declare(Var &v, Type t, bool is_const, bool is_global, Value val) {
  auto it = vars_.find(v.name);
  if (it != vars_.end()) {
    const ScopeEntry &e = scope_[it->second];
    // a local may take a global's name; nothing else may be redeclared
    if (!e.is_global || is_global) {
      err_.error("Cannot redeclare variable \"" + v.name + "\"", v.loc);
      return;
    }
  }
  v.id = ++next_id_;
  vars_.insert({v.name, v.id});  // keeps an existing (global) entry
  scope_[v.id] = ScopeEntry{t, is_const, is_global, val};
}

bool Analysis::promote(ExprP &e, const Type &t) {
  if (e->type == t) return true;
  if (e->type.is_int() && t.is_float()) {
    if (e->kind == EK::IntLit) {
      auto f = std::make_unique<Expr>(EK::FloatLit, e->loc);
      f->fval = (double) e->ival;
      f->type = Type(TI::Float);
      e = std::move(f);
    } else {
      auto c = std::make_unique<Expr>(EK::Call, e->loc);
      c->name = "float";
      c->call = CallKind::Ctor;
      c->type = Type(TI::Float);
      c->a.push_back(std::move(e));
      e = std::move(c);
    }
    return true;
  }
  return false;
}

Value Analysis::eval(const Expr &e) {
  switch (e.kind) {
    case EK::BoolLit: return Value::of_bool(e.bval);
    case EK::IntLit: return Value::of_int(e.ival);
    case EK::FloatLit: return Value::of_float(e.fval);
    case EK::StringLit: return Value::of_string(e.sval);
    case EK::Var: {
      auto it = scope_.find(e.var.id);
      return it == scope_.end() ? Value() : it->second.val;
    }
    case EK::Call: {
      if (e.call == CallKind::Ctor) {
        switch (e.type.id) {
          case TI::Bool: return eval(*e.a[0]).to_bool();
          case TI::Int: return eval(*e.a[0]).to_int();
          case TI::Float: return eval(*e.a[0]).to_float();
          case TI::Vec2:
          case TI::Vec3: {
            std::vector<double> c;
            for (const auto &x : e.a) {
              Value v = eval(*x).to_float_implicit();
              if (v.invalid()) return Value();
              c.push_back(v.f());
            }
            const size_t n = e.type.id == TI::Vec2 ? 2 : 3;
            if (c.size() == 1) c.assign(n, c[0]);
            if (c.size() != n) return Value();
            return n == 2 ? Value::of_vec2(c[0], c[1]) : Value::of_vec3(c[0], c[1], c[2]);
          }
          default: return Value();
        }
      }
      if (e.call == CallKind::Builtin) {
        std::vector<Value> args;
        for (const auto &x : e.a) {
          Value v = eval(*x);
          if (v.invalid()) return Value();
          args.push_back(v);
        }
        return fold_builtin(e.sig, args);
      }
      return Value();
    }
    case EK::Unary: {
      Value v = eval(*e.a[0]);
      return v.invalid() ? Value() : fold_unary(e.uop, v);
    }
    case EK::Binary: {
      Value l = eval(*e.a[0]), r = eval(*e.a[1]);
      if (l.invalid() || r.invalid()) return Value();
      return fold_binary(e.bop, l, r);
    }
    default:
      return Value();
  }
}

// ---- declarations --------------------------------------------------------------

void Analysis::agent(AgentDecl &a) {
  if (agents_.count(a.name)) {
    err_.error("Redefinition of agent " + a.name, a.loc);
  } else {
    agents_[a.name] = &a;
    a.index = (int) script_.agents.size();
    script_.agents.push_back(&a);
    // the agent type as a value, for count(Agent)
    const int id = ++next_id_;
    vars_.insert({a.name, id});
    scope_[id] = ScopeEntry{Type::agent_of(TI::AgentType, &a), true, true, Value()};
  }
  for (auto &m : a.members) m->type.resolved = resolve(m->type);
}

void Analysis::func(FuncDecl &f) {
  push();
  bool entered = true;
  do {
    Type ret = resolve(f.ret);
    if (ret.invalid()) {
      entered = false;
      break;
    }
    std::vector<Type> ptypes;
    bool bad = false;
    for (auto &p : f.params) {
      Type t = resolve(p->type);
      if (t.invalid()) {
        bad = true;
        break;
      }
      ptypes.push_back(t);
      if (!f.is_step() && p->out)
        err_.error("Out variable (-> " + p->out->name + ") can only be used in step functions",
                   p->out->loc);
    }
    if (bad) {
      entered = false;
      break;
    }
    auto it = funcs_.byname.find(f.name);
    if (it != funcs_.byname.end()) {
      for (auto &e : it->second) {
        if (conflicting(e.sig, ptypes)) {
          err_.error("Declaration " + f.name + args_text(ptypes) +
                         " conflicts with previous declaration " + f.name +
                         args_text(e.sig.params),
                     f.loc);
          entered = false;
          break;
        }
      }
      if (!entered) break;
    }
    std::string name = f.name;
    if (it != funcs_.byname.end()) name += "_" + std::to_string(it->second.size());
    // getColor()/getSize() exist for the Java backends' visualisation only
    if ((f.name == "getColor" || f.name == "getSize") && backend_ != "mason" &&
        backend_ != "dmason") {
      cur_ = &f;
      break;
    }
    f.sig = sig(f.name, name, ptypes, ret);
    f.sig.decl = &f;
    f.registered = true;
    funcs_.add(f.sig);
    func_decls_[f.name] = &f;
    cur_ = &f;
    script_.funcs.push_back(&f);
    if (f.is_main()) script_.main_func = &f;
  } while (false);
  (void) entered;
  // children: return type, parameters, body
  f.ret.resolved = resolve(f.ret);
  for (auto &p : f.params) {
    p->type.resolved = resolve(p->type);
    declare(p->var, p->type.resolved, p->out != nullptr, false, Value());
    if (p->out) declare(*p->out, p->type.resolved, false, false, Value());
  }
  for (auto &s : f.body) stmt(*s);
  pop();
  if (f.is_step()) {
    const bool valid = f.params.size() == 1 && f.params[0]->out &&
                       f.params[0]->type.resolved.is_agent();
    if (!valid) err_.error("Step function " + f.name + " does not have a valid signature", f.loc);
  }
}

void Analysis::cnst(ConstDecl &c) {
  script_.consts.push_back(&c);
  c.type.resolved = resolve(c.type);
  expr(c.expr);
  Value val;
  if (c.is_array) {
    const Type elem = c.type.resolved;
    c.type.resolved = Type::array_of(elem);
    if (c.expr->kind != EK::ArrayInit) {
      err_.error("Array must be initialized using array initializer", c.expr->loc);
      return;
    }
    for (auto &x : c.expr->a) {
      if (!promote(x, elem)) {
        err_.error("Element of type " + str(x->type) + " inside initializer for array of type " +
                       str(elem),
                   x->loc);
        return;
      }
    }
    c.expr->type = Type::array_of(elem);
    if (!is_constant_expr(*c.expr)) {
      err_.error("Initializer of global constant must be a constant expression", c.expr->loc);
      return;
    }
  } else {
    val = eval(*c.expr);
    if (val.invalid()) {
      err_.error("Initializer of global constant must be a constant expression", c.expr->loc);
      return;
    }
    c.expr = value_expr(val);
  }
  const Type dt = c.type.resolved;
  if (!promote(c.expr, dt)) {
    err_.error("Trying to assign value of type " + str(c.expr->type) + " to global of type " +
                   str(dt),
               c.expr->loc);
    return;
  }
  auto pit = params_.find(c.var.name);
  if (pit != params_.end()) {
    if (!c.is_param) {
      err_.error("Only constants marked with \"param\" can be specified as parameters", c.var.loc);
      return;
    }
    val = Value::parse(pit->second);
    if (val.invalid()) {
      err_.error("Value \"" + pit->second + "\" provided for parameter \"" + c.var.name +
                     "\" could not parsed",
                 c.var.loc);
      return;
    }
    if (!val.type().promotable_to(dt)) {
      err_.error("Provided parameter \"" + c.var.name + "\" is of type " + str(val.type()) +
                     ", but " + str(dt) + " expected",
                 c.var.loc);
      return;
    }
    c.expr = value_expr(val);
  }
  if (c.is_param) script_.params.insert(c.var.name);
  declare(c.var, c.type.resolved, true, true, val);
}

void Analysis::env(EnvDecl &e) {
  for (auto &m : e.inits) expr(m.expr);
  if (script_.env) {
    err_.error("Script can only contain a single environment specification", e.loc);
    return;
  }
  for (auto &m : e.inits) {
    const Expr &x = *m.expr;
    Value v = eval(x);
    if (v.invalid()) {
      err_.error("Environment member \"" + m.name + "\" must be a constant expression", x.loc);
      return;
    }
    if (m.name == "min") {
      if (!x.type.is_vec()) {
        err_.error("Environment min bound must be float2 or float3", x.loc);
        return;
      }
      e.min = v;
    } else if (m.name == "max") {
      if (!x.type.is_vec()) {
        err_.error("Environment max bound must be float2 or float3", x.loc);
        return;
      }
      e.max = v;
    } else if (m.name == "granularity") {
      if (!x.type.is_num()) {
        err_.error("Environment granularity must be a number", x.loc);
        return;
      }
      e.granularity = v;
    } else {
      err_.error("Unknown environment member \"" + m.name + "\"", m.loc);
      return;
    }
  }
  if (e.max.valid()) {
    e.dim = e.max.is_vec2() ? 2 : 3;
    if (e.min.valid()) {
      if (e.max.type() != e.min.type()) {
        err_.error("min and max environment bounds must have the same type", e.loc);
        return;
      }
    } else {
      e.min = e.dim == 2 ? Value::of_vec2(0, 0) : Value::of_vec3(0, 0, 0);
    }
    e.size = fold_binary(BinOp::Sub, e.max, e.min);
    for (double c : e.size.comps()) {
      if (c < 0) {
        err_.error("Environment minimum should be smaller or equal than the maximum", e.loc);
        return;
      }
    }
  }
  script_.env = &e;
}

// ---- statements ------------------------------------------------------------------

void Analysis::stmt(Stmt &s) {
  switch (s.kind) {
    case SK::Expr:
      expr(s.e0);
      break;
    case SK::Assign: {
      expr(s.e0);
      expr(s.e1);
      // only variables and their members can be assigned
      const Expr *root = s.e0.get();
      while (root->kind == EK::Member) root = root->a[0].get();
      if (root->kind == EK::Var) {
        auto it = scope_.find(root->var.id);
        if (it != scope_.end() && it->second.is_const) {
          err_.error("Trying to assign to immutable variable", s.e0->loc);
          break;
        }
      }
      const Type l = s.e0->type, r = s.e1->type;
      if (l.invalid() || r.invalid()) break;
      if (!r.promotable_to(l))
        err_.error("Cannot assign value of type " + str(r) + " to variable of type " + str(l),
                   s.e1->loc);
      break;
    }
    case SK::AssignOp: {
      expr(s.e0);
      expr(s.e1);
      const Expr *root = s.e0.get();
      while (root->kind == EK::Member) root = root->a[0].get();
      if (root->kind == EK::Var) {
        auto it = scope_.find(root->var.id);
        if (it != scope_.end() && it->second.is_const)
          err_.error("Trying to assign to immutable variable", s.e0->loc);
      }
      break;
    }
    case SK::Block:
      push();
      for (auto &x : s.body) stmt(*x);
      pop();
      break;
    case SK::VarDecl: {
      s.type.resolved = resolve(s.type);
      if (s.e0) expr(s.e0);
      declare(s.var, s.type.resolved, false, false, Value());
      if (!s.e0) {
        err_.error("Variable declaration must have an initializer", s.loc);
        break;
      }
      const Type dt = s.type.resolved, it = s.e0->type;
      if (dt.invalid() || it.invalid()) break;
      if (!promote(s.e0, dt))
        err_.error("Trying to assign value of type " + str(it) + " to variable of type " + str(dt),
                   s.e0->loc);
      break;
    }
    case SK::If: {
      expr(s.e0);
      stmt(*s.s0);
      if (s.s1) stmt(*s.s1);
      const Type t = s.e0->type;
      if (t.invalid()) break;
      if (!t.is_bool())
        err_.error("if() condition must be bool, but received " + str(t), s.e0->loc);
      break;
    }
    case SK::While: {
      loops_++;
      expr(s.e0);
      stmt(*s.s0);
      loops_--;
      const Type t = s.e0->type;
      if (t.invalid()) break;
      if (!t.is_bool())
        err_.error("while() condition must be bool, but received " + str(t), s.e0->loc);
      break;
    }
    case SK::For: {
      loops_++;
      push();
      const Type dt = resolve(s.type);
      declare(s.var, dt, true, false, Value());
      const bool is_near = s.e0->kind == EK::Call && s.e0->name == "near";
      if (is_near) {
        if (!dt.is_agent()) {
          err_.error("Type specified in for-near loop is not an agent", s.type.loc);
        } else if (cur_ && cur_->near_agent) {
          err_.error("Multiple for-near loops in a single step function", s.loc);
        } else if (!dt.agent->position()) {
          err_.error("Cannot use for-near loop on agent without position member", s.loc);
        } else {
          if (cur_) cur_->near_agent = dt.agent;
          s.for_kind = ForKind::Near;
          collect_var_ = s.var.id;
        }
      }
      s.type.resolved = resolve(s.type);
      expr(s.e0);
      stmt(*s.s0);
      loops_--;
      pop();
      if (s.for_kind == ForKind::Near) {
        collect_var_ = 0;
        radii_.push_back(s.e0->a.size() > 1 ? eval(*s.e0->a[1]) : Value());
        break;
      }
      const Type et = s.e0->type;
      if (!et.is_array()) {
        err_.error("Can only use for with array type, received " + str(et), s.e0->loc);
        break;
      }
      if (!et.elem().compatible_with(s.type.resolved)) {
        err_.error("For expression type " + str(et) + " not compatible with declared " +
                       str(s.type.resolved),
                   s.e0->loc);
        break;
      }
      s.for_kind = s.e0->kind == EK::Binary && s.e0->bop == BinOp::Range ? ForKind::Range
                                                                          : ForKind::Normal;
      break;
    }
    case SK::Simulate: {
      expr(s.e0);
      if (script_.simulate) {
        err_.error("Script can only contain a single simulate statement", s.loc);
        break;
      }
      if (!s.e0->type.is_int()) {
        err_.error("Number of timesteps must be an integer, " + str(s.e0->type) + " given",
                   s.e0->loc);
        break;
      }
      bool ok = true;
      for (const std::string &n : s.steps) {
        auto it = func_decls_.find(n);
        if (it == func_decls_.end()) {
          err_.error("Unknown step function \"" + n + "\"", s.loc);
          ok = false;
          break;
        }
        FuncDecl &f = *it->second;
        if (!f.is_any_step()) {
          err_.error("Function \"" + n + "\" is not a step function", s.loc);
          ok = false;
          break;
        }
        if (f.is_seq_step()) {
          if (s.seq_decl) {
            err_.error("Can only use single sequential step function", s.loc);
            ok = false;
            break;
          }
          s.seq_decl = &f;
        } else {
          if (s.seq_decl) {
            err_.error("Sequential step function must be last", s.loc);
            ok = false;
            break;
          }
          s.step_decls.push_back(&f);
        }
      }
      if (ok) script_.simulate = &s;
      break;
    }
    case SK::Return: {
      if (s.e0) expr(s.e0);
      if (!cur_) break;
      const Type rt = cur_->ret.resolved;
      if (rt.is_void()) {
        if (s.e0) err_.error("Cannot return value from void function", s.e0->loc);
      } else if (!s.e0) {
        err_.error("Return from non-void function must specify value", s.loc);
      } else if (!promote(s.e0, rt)) {
        err_.error("Trying to return " + str(s.e0->type) + " from function with return type " +
                       str(rt),
                   s.e0->loc);
      }
      break;
    }
    case SK::Break:
      if (loops_ == 0) err_.error("Cannot use break outside a loop", s.loc);
      break;
    case SK::Continue:
      if (loops_ == 0) err_.error("Cannot use continue outside a loop", s.loc);
      break;
  }
}

// ---- expressions -------------------------------------------------------------------

void Analysis::expr(ExprP &ep) {
  Expr &e = *ep;
  auto prom = [this](ExprP &x, const Type &t) { return promote(x, t); };
  switch (e.kind) {
    case EK::Var: {
      auto it = vars_.find(e.var.name);
      if (it == vars_.end()) {
        err_.error("Use of undeclared variable " + e.var.name, e.loc);
        return;
      }
      e.var.id = it->second;
      e.type = scope_[e.var.id].type;
      return;
    }
    case EK::IntLit: e.type = Type(TI::Int); return;
    case EK::FloatLit: e.type = Type(TI::Float); return;
    case EK::BoolLit: e.type = Type(TI::Bool); return;
    case EK::StringLit: e.type = Type(TI::String); return;
    case EK::Unary: {
      expr(e.a[0]);
      const Type t = e.a[0]->type;
      if (t.invalid()) return;
      Type r;
      switch (e.uop) {
        case UnOp::Plus:
        case UnOp::Minus: r = t.is_num_or_vec() ? t : Type(); break;
        case UnOp::Not: r = t.is_bool() ? t : Type(); break;
        case UnOp::BitNot: r = t.is_int() ? t : Type(); break;
      }
      e.type = r;
      if (r.invalid())
        err_.error(std::string("Type mismatch: Applying unary operator \"") + sigil(e.uop) +
                       "\" to " + str(t),
                   e.loc);
      return;
    }
    case EK::Binary: {
      expr(e.a[0]);
      expr(e.a[1]);
      if (e.a[0]->type.invalid() || e.a[1]->type.invalid()) return;
      ExprP &L = e.a[0], &R = e.a[1];
      auto pb = [&] { return prom(L, R->type) || prom(R, L->type); };
      const Type l = L->type, r = R->type;
      Type t;
      switch (e.bop) {
        case BinOp::Add:
        case BinOp::Sub:
          if (!l.is_num_or_vec() || !r.is_num_or_vec()) break;
          if (l != r) {
            if (pb()) t = Type(TI::Float);
            break;
          }
          t = l;
          break;
        case BinOp::Mul:
        case BinOp::Div:
          if (!l.is_num_or_vec() || !r.is_num_or_vec()) break;
          if (l.is_vec() && r.is_vec()) break;
          if (l.is_vec()) {
            prom(R, Type(TI::Float));
            t = l;
            break;
          }
          if (r.is_vec()) {
            if (e.bop == BinOp::Div) break;
            prom(L, Type(TI::Float));
            t = r;
            break;
          }
          if (l != r) {
            if (pb()) t = Type(TI::Float);
            break;
          }
          t = l;
          break;
        case BinOp::Mod:
          if (l.is_int() && r.is_int()) t = Type(TI::Int);
          else if (pb()) t = Type(TI::Float);
          break;
        case BinOp::BitOr:
        case BinOp::BitAnd:
        case BinOp::BitXor:
        case BinOp::Shl:
        case BinOp::Shr:
          if (l.is_int() && r.is_int()) t = Type(TI::Int);
          break;
        case BinOp::Eq:
        case BinOp::Ne:
          if ((l.is_num() && r.is_num()) || (l.is_vec() && l == r)) t = Type(TI::Bool);
          break;
        case BinOp::Lt:
        case BinOp::Le:
        case BinOp::Gt:
        case BinOp::Ge:
          if (l.is_num() && r.is_num()) t = Type(TI::Bool);
          break;
        case BinOp::And:
        case BinOp::Or:
          if (l.is_bool() && r.is_bool()) t = Type(TI::Bool);
          break;
        case BinOp::Range:
          if (l.is_int() && r.is_int()) t = Type::array_of(Type(TI::Int));
          break;
      }
      e.type = t;
      if (t.invalid()) {
        err_.error("Type mismatch (" + str(L->type) + " " + sigil(e.bop) + " " + str(R->type) + ")",
                   e.loc);
        return;
      }
      // scalar * vector is written vector * scalar
      if (e.bop == BinOp::Mul && R->type.is_vec()) std::swap(e.a[0], e.a[1]);
      return;
    }
    case EK::Call:
      for (auto &x : e.a) expr(x);
      call(e);
      return;
    case EK::Member:
      expr(e.a[0]);
      member(e);
      return;
    case EK::EnvAccess: {
      if (!script_.env) {
        err_.error("Cannot access environment prior to its declaration", e.loc);
        return;
      }
      if (e.name == "max") ep = value_expr(script_.env->max);
      else if (e.name == "min") ep = value_expr(script_.env->min);
      else err_.error("Unknown environment member \"" + e.name + "\"", e.loc);
      return;
    }
    case EK::Index: {
      expr(e.a[0]);
      expr(e.a[1]);
      if (!e.a[0]->type.is_array()) {
        err_.error("Can only index into arrays", e.a[0]->loc);
        return;
      }
      if (!e.a[1]->type.is_int()) {
        err_.error("Array offset must be an integer", e.a[1]->loc);
        return;
      }
      e.type = e.a[0]->type.elem();
      return;
    }
    case EK::Ternary: {
      expr(e.a[0]);
      expr(e.a[1]);
      expr(e.a[2]);
      const Type it = e.a[1]->type, et = e.a[2]->type;
      if (it.invalid() || et.invalid()) return;
      if (prom(e.a[1], et)) e.type = et;
      else if (prom(e.a[2], it)) e.type = it;
      else
        err_.error("Branches of ternary operator have divergent types " + str(it) + " and " +
                       str(et),
                   e.loc);
      return;
    }
    case EK::AgentNew: {
      for (auto &m : e.inits) expr(m.expr);
      auto it = agents_.find(e.name);
      if (it == agents_.end()) {
        err_.error("Unknown agent type \"" + e.name + "\"", e.loc);
        return;
      }
      AgentDecl &ag = *it->second;
      std::set<std::string> seen;
      for (auto &m : e.inits) {
        AgentMemberDecl *md = find_member(ag, m.name);
        if (!md) {
          err_.error("Agent has no member \"" + m.name + "\"", m.loc);
          return;
        }
        const Type mt = md->type.resolved, xt = m.expr->type;
        if (xt.invalid()) return;
        if (!prom(m.expr, mt)) {
          err_.error("Trying to initialize member of type " + str(mt) + " from expression of type " +
                         str(xt),
                     m.expr->loc);
          return;
        }
        seen.insert(m.name);
      }
      for (auto &md : ag.members) {
        if (!seen.count(md->name)) {
          err_.error("Agent member \"" + md->name + "\" has not been initialized", e.loc);
          return;
        }
      }
      e.type = Type::agent_of(TI::Agent, &ag);
      return;
    }
    case EK::ArrayInit:
      for (auto &x : e.a) expr(x);
      return;
    case EK::NewArray: {
      const Type et = resolve(e.elem);
      e.elem.resolved = et;
      expr(e.a[0]);
      if (et.is_array()) {
        err_.error("Cannot instantiate nested array type", e.elem.loc);
        return;
      }
      e.type = Type::array_of(et);
      return;
    }
  }
}

void Analysis::member(Expr &e) {
  const Type t = e.a[0]->type;
  if (t.is_vec()) {
    if (!(e.name == "x" || e.name == "y" || (t.id == TI::Vec3 && e.name == "z"))) {
      err_.error("Vector has no member \"" + e.name + "\"", e.loc);
      return;
    }
    e.type = Type(TI::Float);
  } else if (t.is_agent() || t.is_agent_type()) {
    AgentMemberDecl *m = t.agent ? find_member(*t.agent, e.name) : nullptr;
    if (!m) {
      err_.error("Agent has no member \"" + e.name + "\"", e.loc);
      return;
    }
    e.type = t.is_agent() ? m->type.resolved : Type::member_of(t.agent, m);
  } else {
    err_.error("Can only access members on agent or vector type", e.loc);
    return;
  }
  const Expr &v = *e.a[0];
  if (v.kind == EK::Var && collect_var_ && v.var.id == collect_var_ && cur_)
    cur_->near_members.insert(e.name);
}

void Analysis::call(Expr &e) {
  std::vector<Type> args;
  for (auto &x : e.a) {
    if (x->type.invalid()) return;
    args.push_back(x->type);
  }
  const Type ct = try_simple(e.name);
  if (!ct.invalid()) {
    if (!ctor_valid(ct, args)) {
      err_.error("Type constructor called with invalid arguments: " + e.name + args_text(args),
                 e.loc);
      return;
    }
    e.call = CallKind::Ctor;
    e.type = ct;
    return;
  }
  auto it = funcs_.byname.find(e.name);
  if (it == funcs_.byname.end()) {
    err_.error("Call to unknown function \"" + e.name + "\"", e.loc);
    return;
  }
  const FunctionTable::Entry *ent = nullptr;
  for (auto &x : it->second) {
    if (compatible(x, args)) {
      ent = &x;
      break;
    }
  }
  if (!ent) {
    std::string m = "Function called with invalid arguments: " + e.name + args_text(args) +
                    ", expected ";
    bool first = true;
    for (auto &x : it->second) {
      m += (first ? "" : " or ") + e.name + args_text(x.sig.params);
      first = false;
    }
    err_.error(m, e.loc);
    return;
  }
  const Signature &s = ent->sig;
  if (s.decl && s.decl->is_any_step()) {
    err_.error("Cannot directly call step function " + e.name + "()", e.loc);
    return;
  }
  if ((s.flags & Signature::SeqStepOnly) && !(cur_ && cur_->is_seq_step())) {
    err_.error(e.name + "() can only be used inside a sequential step function", e.loc);
    return;
  }
  if (e.name == "removeCurrent") {
    if (!(cur_ && cur_->is_step())) {
      err_.error("removeCurrent() can only be used inside a step function", e.loc);
      return;
    }
    cur_->uses_removal = true;
    if (cur_->step_agent()) cur_->step_agent()->uses_removal = true;
    script_.uses_removal = true;
  }
  if (e.name == "add") {
    if (!(cur_ && (cur_->is_step() || cur_->is_main()))) {
      err_.error("add() can only be used in main() or a step function", e.loc);
      return;
    }
    if (cur_->is_step()) {
      const Expr &arg = *e.a[0];
      if (arg.kind != EK::AgentNew) {
        err_.error("Argument of add() must be an agent creation expression", e.loc);
        return;
      }
      auto ai = agents_.find(arg.name);
      if (ai == agents_.end()) return;
      AgentDecl *added = ai->second;
      if (AgentMemberDecl *pm = added->position()) {
        for (auto &mi : arg.inits)
          if (mi.name == pm->name && !is_in_pos(*mi.expr, pm->name, *cur_))
            script_.addition_at_other_pos = true;
      }
      if (cur_->added_agent) {
        err_.error("Only one agent per step function may be added at runtime", e.loc);
        return;
      }
      cur_->added_agent = added;
      script_.uses_addition = true;
    }
  }
  e.call = s.decl ? CallKind::User : CallKind::Builtin;
  Signature c = concrete(*ent, args);
  e.sig = c.name;
  e.sig_params = c.params;
  e.type = c.ret;
  e.callee = s.decl;
  if (e.name == "log_csv") script_.uses_logging = true;
  if (e.name == "getLastExecTime") script_.uses_timing = true;
  if (cur_ && (e.name == "random" || e.name == "randomInt")) cur_->uses_rng = true;
  if (cur_ && s.decl && s.decl->uses_rng) cur_->uses_rng = true;
}

void Analysis::finish() {
  EnvDecl *env = script_.env;
  for (AgentDecl *a : script_.agents) {
    AgentMemberDecl *m = a->position();
    if (!m) continue;
    if (!env || env->dim < 0) {
      err_.error("An environment { } declaration is required to use position members", m->loc);
      return;
    }
    const Type t = m->type.resolved;
    if (!t.is_vec() || t.vec_len() != env->dim) {
      err_.error("Dimensionality of position member does not match environment dimension", m->loc);
      return;
    }
  }
  if (env && env->granularity.invalid()) {
    // the largest constant near radius
    Value best;
    for (const Value &r : radii_) {
      if (r.invalid()) continue;
      if (best.invalid() || r.as_float() > best.as_float()) best = r;
    }
    if (best.invalid()) {
      err_.error("Could not automatically determine partitioning granularity. Please explicitly "
                 "specify it in the environment { } declaration",
                 env->loc);
      return;
    }
    env->granularity = best;
  }
  if (!script_.main_func) {
    err_.error("Script must have a main function", Loc{});
    return;
  }
  if (script_.simulate) {
    bool top = false;
    for (auto &s : script_.main_func->body)
      if (s.get() == script_.simulate) top = true;
    if (!top) {
      err_.error("Simulate statement cannot be used conditionally", script_.simulate->loc);
      return;
    }
  }
  for (const auto &kv : params_) {
    if (!script_.params.count(kv.first)) {
      err_.error("Unknown parameter \"" + kv.first + "\" specified through -P", Loc{});
      return;
    }
  }
}

}  // namespace ablc
