// types.hpp -- types of the .abl language and compile-time values.
//
// Type: the language's value types (src/Type.hpp:28-183 of the reference
// defines the same set): bool, int, float, float2, float3, string, agents,
// arrays of them, and the two reduction-argument types (an agent type, one of
// its members).  Value: a compile-time constant, the result of folding global
// initialisers, environment bounds, near radii and -P overrides
// (src/Value.cpp:133-322 gives the folding rules this follows: int
// arithmetic stays int, mixed arithmetic goes through double, vectors are
// component-wise, and the six one-argument math builtins plus pow fold).
#pragma once

#include <ostream>
#include <string>
#include <vector>

namespace ablc {

struct AgentDecl;
struct AgentMemberDecl;

enum class TI {
  Invalid, Void, Bool, Int, Float, String, Vec2, Vec3, Agent, Array, AgentType, AgentMember,
  Unresolved,
};

struct Type {
  TI id = TI::Invalid;
  TI base = TI::Invalid;             // Array element type
  AgentDecl *agent = nullptr;        // Agent, AgentType, AgentMember; Array of agents
  AgentMemberDecl *member = nullptr; // AgentMember

  Type() = default;
  Type(TI t) : id(t) {}
  static Type agent_of(TI t, AgentDecl *a) {
    Type r(t);
    r.agent = a;
    return r;
  }
  static Type member_of(AgentDecl *a, AgentMemberDecl *m) {
    Type r(TI::AgentMember);
    r.agent = a;
    r.member = m;
    return r;
  }
  static Type array_of(const Type &elem) {
    Type r(TI::Array);
    r.base = elem.id;
    r.agent = elem.agent;
    r.member = elem.member;
    return r;
  }
  Type elem() const {
    Type r(base);
    if (base == TI::Agent) r.agent = agent;
    return r;
  }

  bool operator==(const Type &o) const {
    if (id != o.id) return false;
    if (id == TI::Array) return elem() == o.elem();
    return agent == o.agent && member == o.member;
  }
  bool operator!=(const Type &o) const { return !(*this == o); }

  bool invalid() const { return id == TI::Invalid; }
  bool is_void() const { return id == TI::Void; }
  bool is_bool() const { return id == TI::Bool; }
  bool is_int() const { return id == TI::Int; }
  bool is_float() const { return id == TI::Float; }
  bool is_string() const { return id == TI::String; }
  bool is_vec() const { return id == TI::Vec2 || id == TI::Vec3; }
  bool is_num() const { return is_int() || is_float(); }
  bool is_num_or_vec() const { return is_num() || is_vec(); }
  bool is_agent() const { return id == TI::Agent; }
  bool is_array() const { return id == TI::Array; }
  bool is_agent_type() const { return id == TI::AgentType; }
  bool is_agent_member() const { return id == TI::AgentMember; }
  int vec_len() const { return id == TI::Vec2 ? 2 : 3; }

  // `this` may be passed where `o` is expected: identical, an int where a
  // float is expected (promote), or an agent where any agent is expected.
  bool compatible(const Type &o, bool promote) const {
    if (id != o.id) return promote && id == TI::Int && o.id == TI::Float;
    if (id == TI::Agent) return agent == o.agent || o.agent == nullptr;
    if (id == TI::Array) return elem().compatible(o.elem(), false);
    return true;
  }
  bool compatible_with(const Type &o) const { return compatible(o, false); }
  bool promotable_to(const Type &o) const { return compatible(o, true); }
};

// Printed as the reference's diagnostics print types (src/Type.cpp:20-74).
std::ostream &operator<<(std::ostream &s, const Type &t);

struct Vec {
  double x = 0, y = 0, z = 0;
};

// A folded constant.  Ints are held as long, floats as double.
class Value {
 public:
  Value() = default;
  static Value of_bool(bool b) { Value v; v.t_ = TI::Bool; v.b_ = b; return v; }
  static Value of_int(long i) { Value v; v.t_ = TI::Int; v.i_ = i; return v; }
  static Value of_float(double f) { Value v; v.t_ = TI::Float; v.f_ = f; return v; }
  static Value of_string(const std::string &s) { Value v; v.t_ = TI::String; v.s_ = s; return v; }
  static Value of_vec2(double x, double y) { Value v; v.t_ = TI::Vec2; v.v_ = {x, y, 0}; return v; }
  static Value of_vec3(double x, double y, double z) {
    Value v; v.t_ = TI::Vec3; v.v_ = {x, y, z}; return v;
  }
  // -P value text: true/false, an int (the whole text), a double, else invalid
  static Value parse(const std::string &text);

  bool valid() const { return t_ != TI::Invalid; }
  bool invalid() const { return t_ == TI::Invalid; }
  Type type() const { return Type(t_); }
  bool is_bool() const { return t_ == TI::Bool; }
  bool is_int() const { return t_ == TI::Int; }
  bool is_float() const { return t_ == TI::Float; }
  bool is_num() const { return is_int() || is_float(); }
  bool is_vec() const { return t_ == TI::Vec2 || t_ == TI::Vec3; }
  bool is_vec2() const { return t_ == TI::Vec2; }
  bool is_vec3() const { return t_ == TI::Vec3; }
  bool b() const { return b_; }
  long i() const { return i_; }
  double f() const { return f_; }
  double as_float() const { return t_ == TI::Int ? (double) i_ : f_; }
  const Vec &v() const { return v_; }
  const std::string &s() const { return s_; }
  std::vector<double> comps() const {
    if (t_ == TI::Vec2) return {v_.x, v_.y};
    if (t_ == TI::Vec3) return {v_.x, v_.y, v_.z};
    return {};
  }

  // Explicit conversions of type constructors bool(x), int(x), float(x).
  Value to_bool() const;
  Value to_int() const;
  Value to_float() const;
  Value to_float_implicit() const;  // int -> float, float stays, else invalid

 private:
  TI t_ = TI::Invalid;
  bool b_ = false;
  long i_ = 0;
  double f_ = 0.0;
  Vec v_;
  std::string s_;
};

}  // namespace ablc
