// fold.cpp -- compile-time values: -P parsing, conversions, folding of
// unary/binary operators and math builtins, and the printing of types in
// diagnostics.
//
// The folding rules are the reference compiler's (src/Value.cpp:57-322),
// since folded constants are what the generated programs compute with:
// int op int stays int (C semantics: truncating division, %), any other
// numeric mix is computed in double, vectors fold component-wise with a
// scalar on either side of *, comparisons of numbers compare as int or as
// double.  One quirk is kept on purpose: equality of two float3 constants
// compares their x and y components only, as the reference's folder does,
// so constant conditions fold the same way.
#include <cmath>
#include <map>
#include <stdexcept>

#include "fold.hpp"

namespace ablc {

const char *sigil(UnOp op) {
  switch (op) {
    case UnOp::Plus: return "+";
    case UnOp::Minus: return "-";
    case UnOp::Not: return "!";
    case UnOp::BitNot: return "~";
  }
  return "?";
}

const char *sigil(BinOp op) {
  switch (op) {
    case BinOp::Add: return "+";
    case BinOp::Sub: return "-";
    case BinOp::Mul: return "*";
    case BinOp::Div: return "/";
    case BinOp::Mod: return "%";
    case BinOp::BitAnd: return "&";
    case BinOp::BitXor: return "^";
    case BinOp::BitOr: return "|";
    case BinOp::Shl: return "<<";
    case BinOp::Shr: return ">>";
    case BinOp::Eq: return "==";
    case BinOp::Ne: return "!=";
    case BinOp::Lt: return "<";
    case BinOp::Le: return "<=";
    case BinOp::Gt: return ">";
    case BinOp::Ge: return ">=";
    case BinOp::And: return "&&";
    case BinOp::Or: return "||";
    case BinOp::Range: return "..";
  }
  return "?";
}

std::ostream &operator<<(std::ostream &s, const Type &t) {
  switch (t.id) {
    case TI::Invalid: return s << "INVALID";
    case TI::Void: return s << "void";
    case TI::Bool: return s << "bool";
    case TI::Int: return s << "int";
    case TI::Float: return s << "float";
    case TI::String: return s << "string";
    case TI::Vec2: return s << "float2";
    case TI::Vec3: return s << "float3";
    case TI::Agent: return s << (t.agent ? t.agent->name : std::string("agent"));
    case TI::Array: return s << t.elem() << "[]";
    case TI::AgentType:
      s << "agentType";
      if (t.agent) s << "{" << t.agent->name << "}";
      return s;
    case TI::AgentMember:
      s << "agentMember";
      if (t.agent) s << "{" << t.agent->name << "." << t.member->name << "}";
      return s;
    case TI::Unresolved: return s << "unresolved";
  }
  return s;
}

Value Value::parse(const std::string &text) {
  if (text == "true") return of_bool(true);
  if (text == "false") return of_bool(false);
  try {
    size_t pos = 0;
    const long v = std::stoi(text, &pos);
    if (pos == text.size()) return of_int(v);
  } catch (const std::logic_error &) {
  }
  try {
    size_t pos = 0;
    const double d = std::stod(text, &pos);
    if (pos == text.size()) return of_float(d);
  } catch (const std::logic_error &) {
  }
  return Value();
}

Value Value::to_bool() const {
  if (is_bool()) return *this;
  if (is_int()) return of_bool(i_ != 0);
  if (is_float()) return of_bool(f_ != 0.0);
  return Value();
}
Value Value::to_int() const {
  if (is_bool()) return of_int(b_ ? 1 : 0);
  if (is_int()) return *this;
  if (is_float()) return of_int((long) f_);
  return Value();
}
Value Value::to_float() const {
  if (is_bool()) return of_float(b_ ? 1.0 : 0.0);
  if (is_int()) return of_float((double) i_);
  if (is_float()) return *this;
  return Value();
}
Value Value::to_float_implicit() const {
  if (is_int()) return of_float((double) i_);
  if (is_float()) return *this;
  return Value();
}

Value fold_unary(UnOp op, const Value &v) {
  switch (op) {
    case UnOp::Plus:
      return v.is_num() || v.is_vec() ? v : Value();
    case UnOp::Minus:
      if (v.is_int()) return Value::of_int(-v.i());
      if (v.is_float()) return Value::of_float(-v.f());
      if (v.is_vec2()) return Value::of_vec2(-v.v().x, -v.v().y);
      if (v.is_vec3()) return Value::of_vec3(-v.v().x, -v.v().y, -v.v().z);
      return Value();
    case UnOp::Not:
      return v.is_bool() ? Value::of_bool(!v.b()) : Value();
    case UnOp::BitNot:
      return v.is_int() ? Value::of_int(~v.i()) : Value();
  }
  return Value();
}

namespace {

int compare_nums(const Value &l, const Value &r) {
  if (l.is_int() && r.is_int()) return l.i() > r.i() ? 1 : l.i() < r.i() ? -1 : 0;
  const double a = l.as_float(), b = r.as_float();
  return a > b ? 1 : a < b ? -1 : 0;
}

Value vec_of(const Value &like, double x, double y, double z) {
  return like.is_vec2() ? Value::of_vec2(x, y) : Value::of_vec3(x, y, z);
}

}  // namespace

Value fold_binary(BinOp op, const Value &l, const Value &r) {
  const bool ii = l.is_int() && r.is_int();
  const bool nn = l.is_num() && r.is_num();
  const bool same_vec = (l.is_vec2() && r.is_vec2()) || (l.is_vec3() && r.is_vec3());
  switch (op) {
    case BinOp::Add:
      if (ii) return Value::of_int(l.i() + r.i());
      if (nn) return Value::of_float(l.as_float() + r.as_float());
      if (same_vec)
        return vec_of(l, l.v().x + r.v().x, l.v().y + r.v().y, l.v().z + r.v().z);
      return Value();
    case BinOp::Sub:
      if (ii) return Value::of_int(l.i() - r.i());
      if (nn) return Value::of_float(l.as_float() - r.as_float());
      if (same_vec)
        return vec_of(l, l.v().x - r.v().x, l.v().y - r.v().y, l.v().z - r.v().z);
      return Value();
    case BinOp::Mul:
      if (ii) return Value::of_int(l.i() * r.i());
      if (nn) return Value::of_float(l.as_float() * r.as_float());
      if (l.is_vec() && r.is_num()) {
        const double f = r.as_float();
        return vec_of(l, l.v().x * f, l.v().y * f, l.v().z * f);
      }
      if (l.is_num() && r.is_vec()) {
        const double f = l.as_float();
        return vec_of(r, r.v().x * f, r.v().y * f, r.v().z * f);
      }
      return Value();
    case BinOp::Div:
      if (ii) return r.i() == 0 ? Value() : Value::of_int(l.i() / r.i());
      if (nn) return Value::of_float(l.as_float() / r.as_float());
      if (l.is_vec() && r.is_num()) {
        const double f = r.as_float();
        return vec_of(l, l.v().x / f, l.v().y / f, l.v().z / f);
      }
      return Value();
    case BinOp::Mod:
      if (ii) return r.i() == 0 ? Value() : Value::of_int(l.i() % r.i());
      if (nn) return Value::of_float(std::fmod(l.as_float(), r.as_float()));
      return Value();
    case BinOp::BitOr: return ii ? Value::of_int(l.i() | r.i()) : Value();
    case BinOp::BitAnd: return ii ? Value::of_int(l.i() & r.i()) : Value();
    case BinOp::BitXor: return ii ? Value::of_int(l.i() ^ r.i()) : Value();
    case BinOp::Shl: return ii ? Value::of_int(l.i() << r.i()) : Value();
    case BinOp::Shr: return ii ? Value::of_int(l.i() >> r.i()) : Value();
    case BinOp::Eq:
    case BinOp::Ne:
      if (same_vec) {
        // x and y only, for float3 too (the reference folder's comparison)
        const bool eq = l.v().x == r.v().x && l.v().y == r.v().y;
        return Value::of_bool(op == BinOp::Eq ? eq : !eq);
      }
      [[fallthrough]];
    case BinOp::Lt:
    case BinOp::Le:
    case BinOp::Gt:
    case BinOp::Ge: {
      if (!nn) return Value();
      const int c = compare_nums(l, r);
      switch (op) {
        case BinOp::Eq: return Value::of_bool(c == 0);
        case BinOp::Ne: return Value::of_bool(c != 0);
        case BinOp::Lt: return Value::of_bool(c < 0);
        case BinOp::Le: return Value::of_bool(c <= 0);
        case BinOp::Gt: return Value::of_bool(c > 0);
        default: return Value::of_bool(c >= 0);
      }
    }
    case BinOp::And:
    case BinOp::Or:
      if (!l.is_bool() || !r.is_bool()) return Value();
      return Value::of_bool(op == BinOp::Or ? (l.b() || r.b()) : (l.b() && r.b()));
    case BinOp::Range:
      return Value();
  }
  return Value();
}

Value fold_builtin(const std::string &sig_name, const std::vector<Value> &args) {
  if (args.size() == 1) {
    if (!args[0].is_num()) return Value();
    static const std::map<std::string, double (*)(double)> kFns = {
        {"sin", ::sin},   {"cos", ::cos},   {"tan", ::tan},   {"sinh", ::sinh}, {"cosh", ::cosh},
        {"tanh", ::tanh}, {"asin", ::asin}, {"acos", ::acos}, {"atan", ::atan}, {"exp", ::exp},
        {"log", ::log},   {"sqrt", ::sqrt}, {"cbrt", ::cbrt}, {"round", ::round},
    };
    auto it = kFns.find(sig_name);
    if (it == kFns.end()) return Value();
    return Value::of_float(it->second(args[0].as_float()));
  }
  if (args.size() == 2) {
    if (sig_name != "pow" || !args[0].is_num() || !args[1].is_num()) return Value();
    return Value::of_float(std::pow(args[0].as_float(), args[1].as_float()));
  }
  return Value();
}

// A literal node for a folded value; vectors become float2/float3
// constructor calls of float literals.
ExprP value_expr(const Value &v) {
  Loc none;
  ExprP e;
  if (v.is_bool()) {
    e = std::make_unique<Expr>(EK::BoolLit, none);
    e->bval = v.b();
  } else if (v.is_int()) {
    e = std::make_unique<Expr>(EK::IntLit, none);
    e->ival = v.i();
  } else if (v.is_float()) {
    e = std::make_unique<Expr>(EK::FloatLit, none);
    e->fval = v.f();
  } else if (v.type().is_string()) {
    e = std::make_unique<Expr>(EK::StringLit, none);
    e->sval = v.s();
  } else if (v.is_vec()) {
    e = std::make_unique<Expr>(EK::Call, none);
    e->name = v.is_vec2() ? "float2" : "float3";
    e->call = CallKind::Ctor;
    for (double c : v.comps()) {
      auto lit = std::make_unique<Expr>(EK::FloatLit, none);
      lit->fval = c;
      lit->type = Type(TI::Float);
      e->a.push_back(std::move(lit));
    }
  } else {
    return nullptr;
  }
  e->type = v.type();
  return e;
}

}  // namespace ablc
