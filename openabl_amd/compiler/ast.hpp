// ast.hpp -- syntax tree of the .abl language, as the hip backend's front end
// builds and annotates it.
//
// The node set follows the language the reference's grammar accepts
// (src/Parser.y:160-364): agents with (position) members, constants and
// params, one environment block, functions, step and sequential-step
// functions, and the statements and expressions inside them.  Nodes are
// "fat": one struct per family (expression, statement) with a kind tag and
// the fields every kind of that family may use, which keeps the analysis and
// the printers to plain switch statements.
#pragma once

#include <memory>
#include <set>
#include <string>
#include <vector>

#include "types.hpp"

namespace ablc {

// Source position of a node: the line its first token begins on, with the
// reference scanner's rule for where a token begins (lexer.cpp).
struct Loc {
  int line = 1;
};

enum class UnOp { Plus, Minus, Not, BitNot };
enum class BinOp {
  Add, Sub, Mul, Div, Mod, BitAnd, BitXor, BitOr, Shl, Shr,
  Eq, Ne, Lt, Le, Gt, Ge, And, Or, Range,
};
const char *sigil(UnOp op);
const char *sigil(BinOp op);

struct FuncDecl;
struct Expr;
struct Stmt;
using ExprP = std::unique_ptr<Expr>;
using StmtP = std::unique_ptr<Stmt>;

// A declared name (variable, parameter, loop variable); `id` is unique per
// declaration and set by the analysis.
struct Var {
  std::string name;
  Loc loc;
  int id = 0;
};

// A type as written (`float2`, `Boid`), resolved by the analysis.
struct TypeRef {
  std::string name;
  Loc loc;
  Type resolved;
};

enum class EK {
  Var, IntLit, FloatLit, BoolLit, StringLit, Unary, Binary, Call, Member, EnvAccess, Index,
  Ternary, AgentNew, ArrayInit, NewArray,
};

enum class CallKind { Unresolved, User, Builtin, Ctor };

struct MemberInit {
  std::string name;
  ExprP expr;
  Loc loc;
};

struct Expr {
  EK kind;
  Loc loc;
  Type type;  // set by the analysis
  // Var: var; Call/AgentNew: name; Member/EnvAccess: name = member
  Var var;
  std::string name;
  long ival = 0;
  double fval = 0.0;
  bool bval = false;
  std::string sval;
  UnOp uop = UnOp::Plus;
  BinOp bop = BinOp::Add;
  // Unary: a[0]; Binary: a[0] op a[1]; Call: args; Member: a[0]; Index:
  // a[0][a[1]]; Ternary: a[0] ? a[1] : a[2]; ArrayInit: elements;
  // NewArray: a[0] = size
  std::vector<ExprP> a;
  std::vector<MemberInit> inits;  // AgentNew
  TypeRef elem;                   // NewArray element type
  // Call, after analysis
  CallKind call = CallKind::Unresolved;
  std::string sig;          // printed name of the called signature (overloads: name_k)
  std::vector<Type> sig_params;
  FuncDecl *callee = nullptr;

  explicit Expr(EK k, Loc l) : kind(k), loc(l) {}
};

enum class SK { Expr, Assign, AssignOp, Block, VarDecl, If, While, For, Simulate, Return, Break, Continue };
enum class ForKind { Unknown, Normal, Range, Near };

struct Stmt {
  SK kind;
  Loc loc;
  // Expr: e0; Assign/AssignOp: e0 = e1; VarDecl: init e0 (may be null);
  // If: cond e0; While: cond e0; For: e0 = iterated expression;
  // Simulate: e0 = timesteps; Return: e0 (may be null)
  ExprP e0, e1;
  BinOp op = BinOp::Add;      // AssignOp
  std::vector<StmtP> body;    // Block
  StmtP s0, s1;               // If: then/else; While/For: body
  TypeRef type;               // VarDecl, For
  Var var;                    // VarDecl, For
  std::vector<std::string> steps;  // Simulate: step function names, in order
  ForKind for_kind = ForKind::Unknown;
  // Simulate, after analysis
  std::vector<FuncDecl *> step_decls;
  FuncDecl *seq_decl = nullptr;

  explicit Stmt(SK k, Loc l) : kind(k), loc(l) {}
};

struct AgentMemberDecl {
  bool is_position = false;
  TypeRef type;
  std::string name;
  Loc loc;
};

struct AgentDecl {
  std::string name;
  std::vector<std::unique_ptr<AgentMemberDecl>> members;
  Loc loc;
  bool uses_removal = false;
  AgentMemberDecl *position() const {
    for (auto &m : members)
      if (m->is_position) return m.get();
    return nullptr;
  }
  int index = -1;  // declaration order among agents (set by the analysis)
};

struct Param {
  TypeRef type;
  Var var;
  std::unique_ptr<Var> out;  // `in -> out` of step functions
  Loc loc;
};

enum class FuncKind { Normal, Step, SeqStep };

struct Signature {
  std::string orig;   // name as called
  std::string name;   // printed name (overloads: orig_k)
  std::vector<Type> params;
  Type ret;
  unsigned flags = 0;
  FuncDecl *decl = nullptr;
  enum : unsigned { MainOnly = 1, StepOnly = 2, SeqStepOnly = 4 };
};

struct FuncDecl {
  TypeRef ret;
  std::string name;
  std::vector<std::unique_ptr<Param>> params;
  std::vector<StmtP> body;
  FuncKind fkind = FuncKind::Normal;
  Loc loc;
  bool from_lib = false;
  // analysis
  Signature sig;
  bool registered = false;            // got a signature (getColor/getSize do not)
  AgentDecl *near_agent = nullptr;    // agent type of its for-near loop
  std::set<std::string> near_members; // members of that agent the loop body reads
  bool uses_removal = false;
  AgentDecl *added_agent = nullptr;
  bool uses_rng = false;

  bool is_main() const { return name == "main"; }
  bool is_step() const { return fkind == FuncKind::Step; }
  bool is_seq_step() const { return fkind == FuncKind::SeqStep; }
  bool is_any_step() const { return is_step() || is_seq_step(); }
  AgentDecl *step_agent() const { return params[0]->type.resolved.agent; }
};

struct ConstDecl {
  TypeRef type;
  Var var;
  ExprP expr;
  bool is_array = false;
  bool is_param = false;
  Loc loc;
  bool from_lib = false;
};

struct EnvDecl {
  std::vector<MemberInit> inits;
  Loc loc;
  // analysis
  Value min, max, size, granularity;
  int dim = -1;
};

struct Decl {
  enum Kind { Agent, Func, Const, Env } kind;
  std::unique_ptr<AgentDecl> agent;
  std::unique_ptr<FuncDecl> func;
  std::unique_ptr<ConstDecl> cnst;
  std::unique_ptr<EnvDecl> env;
};

// Reduction kinds the sequential step uses (count(Type), count(T.m, v),
// sum(T.m)).
enum class Reduction { CountType, CountMember, SumMember };

struct Script {
  std::vector<Decl> decls;
  // analysis (of the main script; the library's functions and constants are
  // added in front, as the reference adds them)
  std::vector<AgentDecl *> agents;
  std::vector<ConstDecl *> consts;
  std::vector<FuncDecl *> funcs;
  std::set<std::string> params;
  Stmt *simulate = nullptr;
  FuncDecl *main_func = nullptr;
  EnvDecl *env = nullptr;
  bool uses_removal = false, uses_addition = false, uses_logging = false, uses_timing = false;
  bool addition_at_other_pos = false;
};

}  // namespace ablc
