// analysis.hpp -- semantic analysis of .abl scripts (analysis.cpp).
#pragma once

#include <functional>
#include <map>
#include <ostream>
#include <string>
#include <vector>

#include "ast.hpp"

namespace ablc {

// Diagnostics, printed as the reference prints them: "<message> on line <n>".
struct Diagnostics {
  std::ostream *out = nullptr;
  int count = 0;
  std::vector<std::string> lines;
  void error(const std::string &msg, Loc loc) {
    std::string s = msg + " on line " + std::to_string(loc.line);
    if (out) *out << s << "\n";
    lines.push_back(s);
    count++;
  }
};

struct ScopeEntry {
  Type type;
  bool is_const = false;
  bool is_global = false;
  Value val;
};

// Signatures visible to calls: the builtins (src/main.cpp:26-131 lists the
// same set) and the declared functions, in declaration order per name.
struct FunctionTable {
  struct Custom {
    std::function<bool(const std::vector<Type> &)> compatible;
    std::function<Signature(const std::vector<Type> &)> concrete;
  };
  struct Entry {
    Signature sig;
    Custom custom;
  };
  std::map<std::string, std::vector<Entry>> byname;
  void add(const Signature &s, Custom c = {}) { byname[s.orig].push_back({s, c}); }
};

void register_builtins(FunctionTable &f);

class Analysis {
 public:
  Analysis(Script &main, const std::map<std::string, std::string> &params, Diagnostics &diag,
           const std::string &backend);
  void library(Script &lib);  // the standard library, before the main script
  void main_script();

  // variable information, by declaration id
  const ScopeEntry *entry(int id) const {
    auto it = scope_.find(id);
    return it == scope_.end() ? nullptr : &it->second;
  }

 private:
  Script &script_;
  const std::map<std::string, std::string> &params_;
  Diagnostics &err_;
  std::string backend_;
  FunctionTable funcs_;
  std::map<int, ScopeEntry> scope_;
  std::map<std::string, int> vars_;
  std::vector<std::map<std::string, int>> var_stack_;
  FuncDecl *cur_ = nullptr;
  std::map<std::string, AgentDecl *> agents_;
  std::map<std::string, FuncDecl *> func_decls_;
  int collect_var_ = 0;
  std::vector<Value> radii_;
  int loops_ = 0;
  bool lib_ = false;
  int next_id_ = 0;

  void run(Script &s);
  Type resolve(TypeRef &t);
  Type try_simple(const std::string &name) const;
  void declare(Var &v, Type t, bool is_const, bool is_global, Value val);
  void push() { var_stack_.push_back(vars_); }
  void pop() {
    vars_ = var_stack_.back();
    var_stack_.pop_back();
  }
  Value eval(const Expr &e);
  bool promote(ExprP &e, const Type &t);

  void agent(AgentDecl &a);
  void func(FuncDecl &f);
  void cnst(ConstDecl &c);
  void env(EnvDecl &e);
  void stmt(Stmt &s);
  void expr(ExprP &e);
  void call(Expr &e);
  void member(Expr &e);
  void finish();
};

}  // namespace ablc
