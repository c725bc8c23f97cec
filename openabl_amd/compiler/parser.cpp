// parser.cpp -- recursive-descent parser of the .abl language.
//
// Accepts the language of the reference grammar (src/Parser.y:160-364),
// with its operator precedences (C's, lowest first: ?:, ||, &&, |, ^, &,
// == !=, < <= > >=, .., << >>, + -, * / %, unary ! ~ + -, then [] and .
// member access) and its one-token decisions: `T x` at the start of a
// statement declares a variable, `name {` in an expression creates an agent,
// `type name (` at the top level declares a function.
//
// A node's line is the line its first token begins on; a node whose rule
// starts with an omitted optional part (an agent member without `position`)
// begins where the token before it ended, as the reference grammar's
// locations do.
#include "parser.hpp"

#include <stdexcept>

#include "lexer.hpp"

namespace ablc {

namespace {

struct ParseError : std::runtime_error {
  int line;
  ParseError(const std::string &m, int l) : std::runtime_error(m), line(l) {}
};

class Parser {
 public:
  explicit Parser(std::vector<Token> toks) : t_(std::move(toks)) {}

  std::unique_ptr<Script> script() {
    auto s = std::make_unique<Script>();
    while (peek().kind != Tok::End) s->decls.push_back(declaration());
    return s;
  }

 private:
  std::vector<Token> t_;
  size_t p_ = 0;

  const Token &peek(size_t k = 0) const {
    const size_t i = p_ + k;
    return i < t_.size() ? t_[i] : t_.back();
  }
  bool at(Tok k) const { return peek().kind == k; }
  Loc here() const { return Loc{peek().line}; }
  // where an omitted leading part begins: the end of the previous token
  Loc after_prev() const { return Loc{p_ > 0 ? t_[p_ - 1].end_line : 1}; }
  const Token &take() {
    const Token &t = t_[p_];
    if (p_ + 1 < t_.size()) p_++;
    return t;
  }
  [[noreturn]] void fail(const std::string &expecting) {
    const Token &t = peek();
    std::string got = t.kind == Tok::Ident ? "IDENTIFIER" : token_name(t.kind);
    if (t.kind == Tok::Error) got = "invalid character \"" + t.text + "\"";
    throw ParseError("syntax error, unexpected " + got + ", expecting " + expecting, t.line);
  }
  const Token &expect(Tok k) {
    if (!at(k)) fail(token_name(k));
    return take();
  }
  bool accept(Tok k) {
    if (!at(k)) return false;
    take();
    return true;
  }

  TypeRef type_ref() {
    TypeRef r;
    r.loc = here();
    r.name = expect(Tok::Ident).text;
    return r;
  }
  Var var() {
    Var v;
    v.loc = here();
    v.name = expect(Tok::Ident).text;
    return v;
  }

  // ---- declarations ----------------------------------------------------------
  Decl declaration() {
    Decl d;
    if (at(Tok::Agent)) {
      d.kind = Decl::Agent;
      d.agent = agent_decl();
    } else if (at(Tok::Environment)) {
      d.kind = Decl::Env;
      d.env = env_decl();
    } else if (at(Tok::Step) || at(Tok::Sequential)) {
      d.kind = Decl::Func;
      d.func = step_decl();
    } else if (at(Tok::Param)) {
      d.kind = Decl::Const;
      d.cnst = const_decl();
    } else if (at(Tok::Ident) && peek(1).kind == Tok::Ident && peek(2).kind == Tok::LParen) {
      d.kind = Decl::Func;
      d.func = func_decl();
    } else if (at(Tok::Ident)) {
      d.kind = Decl::Const;
      d.cnst = const_decl();
    } else {
      fail("agent, environment, step, sequential, param or a type");
    }
    return d;
  }

  std::unique_ptr<AgentDecl> agent_decl() {
    auto a = std::make_unique<AgentDecl>();
    a->loc = here();
    expect(Tok::Agent);
    a->name = expect(Tok::Ident).text;
    expect(Tok::LBrace);
    while (!at(Tok::RBrace)) {
      auto m = std::make_unique<AgentMemberDecl>();
      if (at(Tok::Position)) {
        m->loc = here();
        take();
        m->is_position = true;
      } else {
        m->loc = after_prev();
      }
      m->type = type_ref();
      m->name = expect(Tok::Ident).text;
      expect(Tok::Semi);
      a->members.push_back(std::move(m));
    }
    expect(Tok::RBrace);
    return a;
  }

  std::vector<MemberInit> member_inits(Tok close) {
    std::vector<MemberInit> v;
    while (!at(close)) {
      MemberInit mi;
      mi.loc = here();
      mi.name = expect(Tok::Ident).text;
      expect(Tok::Colon);
      mi.expr = expression();
      v.push_back(std::move(mi));
      if (!accept(Tok::Comma)) break;
    }
    return v;
  }

  std::unique_ptr<EnvDecl> env_decl() {
    auto e = std::make_unique<EnvDecl>();
    e->loc = here();
    expect(Tok::Environment);
    expect(Tok::LBrace);
    e->inits = member_inits(Tok::RBrace);
    expect(Tok::RBrace);
    return e;
  }

  void params_and_body(FuncDecl &f) {
    expect(Tok::LParen);
    if (!at(Tok::RParen)) {
      for (;;) {
        auto pr = std::make_unique<Param>();
        pr->loc = here();
        pr->type = type_ref();
        pr->var = var();
        if (accept(Tok::Arrow)) pr->out = std::make_unique<Var>(var());
        f.params.push_back(std::move(pr));
        if (!accept(Tok::Comma)) break;
      }
    }
    expect(Tok::RParen);
    expect(Tok::LBrace);
    while (!at(Tok::RBrace)) f.body.push_back(statement());
    expect(Tok::RBrace);
  }

  std::unique_ptr<FuncDecl> step_decl() {
    auto f = std::make_unique<FuncDecl>();
    f->loc = here();
    if (accept(Tok::Sequential)) f->fkind = FuncKind::SeqStep;
    else f->fkind = FuncKind::Step;
    expect(Tok::Step);
    f->ret.name = "void";
    f->name = expect(Tok::Ident).text;
    params_and_body(*f);
    return f;
  }

  std::unique_ptr<FuncDecl> func_decl() {
    auto f = std::make_unique<FuncDecl>();
    f->loc = here();
    f->ret = type_ref();
    f->name = expect(Tok::Ident).text;
    params_and_body(*f);
    return f;
  }

  std::unique_ptr<ConstDecl> const_decl() {
    auto c = std::make_unique<ConstDecl>();
    c->loc = here();
    if (accept(Tok::Param)) c->is_param = true;
    c->type = type_ref();
    c->var = var();
    if (accept(Tok::LBracket)) {
      expect(Tok::RBracket);
      c->is_array = true;
    }
    expect(Tok::Assign);
    if (at(Tok::LBrace)) {
      auto init = std::make_unique<Expr>(EK::ArrayInit, here());
      take();
      for (;;) {
        init->a.push_back(expression());
        if (!accept(Tok::Comma)) break;
        if (at(Tok::RBrace)) break;
      }
      expect(Tok::RBrace);
      c->expr = std::move(init);
    } else {
      c->expr = expression();
    }
    expect(Tok::Semi);
    return c;
  }

  // ---- statements ------------------------------------------------------------
  StmtP statement() {
    const Loc loc = here();
    switch (peek().kind) {
      case Tok::LBrace: {
        auto s = std::make_unique<Stmt>(SK::Block, loc);
        take();
        while (!at(Tok::RBrace)) s->body.push_back(statement());
        take();
        return s;
      }
      case Tok::If: {
        auto s = std::make_unique<Stmt>(SK::If, loc);
        take();
        expect(Tok::LParen);
        s->e0 = expression();
        expect(Tok::RParen);
        s->s0 = statement();
        if (accept(Tok::Else)) s->s1 = statement();
        return s;
      }
      case Tok::While: {
        auto s = std::make_unique<Stmt>(SK::While, loc);
        take();
        expect(Tok::LParen);
        s->e0 = expression();
        expect(Tok::RParen);
        s->s0 = statement();
        return s;
      }
      case Tok::For: {
        auto s = std::make_unique<Stmt>(SK::For, loc);
        take();
        expect(Tok::LParen);
        s->type = type_ref();
        s->var = var();
        expect(Tok::Colon);
        s->e0 = expression();
        expect(Tok::RParen);
        s->s0 = statement();
        return s;
      }
      case Tok::Return: {
        auto s = std::make_unique<Stmt>(SK::Return, loc);
        take();
        if (!at(Tok::Semi)) s->e0 = expression();
        expect(Tok::Semi);
        return s;
      }
      case Tok::Break:
      case Tok::Continue: {
        auto s = std::make_unique<Stmt>(at(Tok::Break) ? SK::Break : SK::Continue, loc);
        take();
        expect(Tok::Semi);
        return s;
      }
      case Tok::Simulate: {
        auto s = std::make_unique<Stmt>(SK::Simulate, loc);
        take();
        expect(Tok::LParen);
        s->e0 = expression();
        expect(Tok::RParen);
        expect(Tok::LBrace);
        s->steps.push_back(expect(Tok::Ident).text);
        while (accept(Tok::Comma)) {
          if (at(Tok::RBrace)) break;
          s->steps.push_back(expect(Tok::Ident).text);
        }
        expect(Tok::RBrace);
        return s;
      }
      default:
        break;
    }
    if (at(Tok::Ident) && peek(1).kind == Tok::Ident) {
      auto s = std::make_unique<Stmt>(SK::VarDecl, loc);
      s->type = type_ref();
      s->var = var();
      if (accept(Tok::Assign)) s->e0 = expression();
      expect(Tok::Semi);
      return s;
    }
    ExprP e = expression();
    static const struct {
      Tok t;
      BinOp op;
    } kAssignOps[] = {
        {Tok::AddAssign, BinOp::Add}, {Tok::SubAssign, BinOp::Sub}, {Tok::MulAssign, BinOp::Mul},
        {Tok::DivAssign, BinOp::Div}, {Tok::ModAssign, BinOp::Mod},
        {Tok::AndAssign, BinOp::BitAnd}, {Tok::XorAssign, BinOp::BitXor},
        {Tok::OrAssign, BinOp::BitOr}, {Tok::ShlAssign, BinOp::Shl}, {Tok::ShrAssign, BinOp::Shr},
    };
    if (accept(Tok::Assign)) {
      auto s = std::make_unique<Stmt>(SK::Assign, loc);
      s->e0 = std::move(e);
      s->e1 = expression();
      expect(Tok::Semi);
      return s;
    }
    for (const auto &ao : kAssignOps) {
      if (accept(ao.t)) {
        auto s = std::make_unique<Stmt>(SK::AssignOp, loc);
        s->op = ao.op;
        s->e0 = std::move(e);
        s->e1 = expression();
        expect(Tok::Semi);
        return s;
      }
    }
    auto s = std::make_unique<Stmt>(SK::Expr, loc);
    s->e0 = std::move(e);
    expect(Tok::Semi);
    return s;
  }

  // ---- expressions -----------------------------------------------------------
  ExprP expression() { return ternary(); }

  ExprP ternary() {
    ExprP c = binary(0);
    if (!at(Tok::Qm)) return c;
    take();
    auto e = std::make_unique<Expr>(EK::Ternary, c->loc);
    e->a.push_back(std::move(c));
    e->a.push_back(expression());
    expect(Tok::Colon);
    e->a.push_back(ternary());
    return e;
  }

  // binary levels, lowest first
  static int level_of(Tok t, BinOp *op) {
    switch (t) {
      case Tok::Or: *op = BinOp::Or; return 0;
      case Tok::And: *op = BinOp::And; return 1;
      case Tok::BitOr: *op = BinOp::BitOr; return 2;
      case Tok::BitXor: *op = BinOp::BitXor; return 3;
      case Tok::BitAnd: *op = BinOp::BitAnd; return 4;
      case Tok::Eq: *op = BinOp::Eq; return 5;
      case Tok::Ne: *op = BinOp::Ne; return 5;
      case Tok::Lt: *op = BinOp::Lt; return 6;
      case Tok::Le: *op = BinOp::Le; return 6;
      case Tok::Gt: *op = BinOp::Gt; return 6;
      case Tok::Ge: *op = BinOp::Ge; return 6;
      case Tok::DotDot: *op = BinOp::Range; return 7;
      case Tok::Shl: *op = BinOp::Shl; return 8;
      case Tok::Shr: *op = BinOp::Shr; return 8;
      case Tok::Add: *op = BinOp::Add; return 9;
      case Tok::Sub: *op = BinOp::Sub; return 9;
      case Tok::Mul: *op = BinOp::Mul; return 10;
      case Tok::Div: *op = BinOp::Div; return 10;
      case Tok::Mod: *op = BinOp::Mod; return 10;
      default: return -1;
    }
  }

  ExprP binary(int min_level) {
    ExprP lhs = unary();
    for (;;) {
      BinOp op;
      const int lv = level_of(peek().kind, &op);
      if (lv < min_level || lv < 0) return lhs;
      take();
      ExprP rhs = binary(lv + 1);
      auto e = std::make_unique<Expr>(EK::Binary, lhs->loc);
      e->bop = op;
      e->a.push_back(std::move(lhs));
      e->a.push_back(std::move(rhs));
      lhs = std::move(e);
      if (op == BinOp::Range && at(Tok::DotDot)) fail("an operator other than ..");
    }
  }

  ExprP unary() {
    UnOp op;
    switch (peek().kind) {
      case Tok::Not: op = UnOp::Not; break;
      case Tok::BitNot: op = UnOp::BitNot; break;
      case Tok::Add: op = UnOp::Plus; break;
      case Tok::Sub: op = UnOp::Minus; break;
      default: return postfix();
    }
    auto e = std::make_unique<Expr>(EK::Unary, here());
    take();
    e->uop = op;
    e->a.push_back(unary());
    return e;
  }

  ExprP postfix() {
    ExprP e = primary();
    for (;;) {
      if (at(Tok::LBracket)) {
        take();
        auto x = std::make_unique<Expr>(EK::Index, e->loc);
        x->a.push_back(std::move(e));
        x->a.push_back(expression());
        expect(Tok::RBracket);
        e = std::move(x);
      } else if (at(Tok::Dot)) {
        take();
        auto x = std::make_unique<Expr>(EK::Member, e->loc);
        x->name = expect(Tok::Ident).text;
        x->a.push_back(std::move(e));
        e = std::move(x);
      } else {
        return e;
      }
    }
  }

  ExprP primary() {
    const Loc loc = here();
    const Token &t = peek();
    switch (t.kind) {
      case Tok::Int: {
        auto e = std::make_unique<Expr>(EK::IntLit, loc);
        e->ival = take().ival;
        return e;
      }
      case Tok::Float: {
        auto e = std::make_unique<Expr>(EK::FloatLit, loc);
        e->fval = take().fval;
        return e;
      }
      case Tok::Bool: {
        auto e = std::make_unique<Expr>(EK::BoolLit, loc);
        e->bval = take().bval;
        return e;
      }
      case Tok::String: {
        auto e = std::make_unique<Expr>(EK::StringLit, loc);
        e->sval = take().text;
        return e;
      }
      case Tok::LParen: {
        take();
        ExprP e = expression();
        expect(Tok::RParen);
        return e;
      }
      case Tok::Environment: {
        take();
        expect(Tok::Dot);
        auto e = std::make_unique<Expr>(EK::EnvAccess, loc);
        e->name = expect(Tok::Ident).text;
        return e;
      }
      case Tok::New: {
        take();
        auto e = std::make_unique<Expr>(EK::NewArray, loc);
        e->elem = type_ref();
        expect(Tok::LBracket);
        e->a.push_back(expression());
        expect(Tok::RBracket);
        return e;
      }
      case Tok::Ident: {
        if (peek(1).kind == Tok::LParen) {
          auto e = std::make_unique<Expr>(EK::Call, loc);
          e->name = take().text;
          take();
          if (!at(Tok::RParen)) {
            for (;;) {
              e->a.push_back(expression());
              if (!accept(Tok::Comma)) break;
              if (at(Tok::RParen)) break;
            }
          }
          expect(Tok::RParen);
          return e;
        }
        if (peek(1).kind == Tok::LBrace) {
          auto e = std::make_unique<Expr>(EK::AgentNew, loc);
          e->name = take().text;
          take();
          e->inits = member_inits(Tok::RBrace);
          expect(Tok::RBrace);
          return e;
        }
        auto e = std::make_unique<Expr>(EK::Var, loc);
        e->var.loc = loc;
        e->var.name = take().text;
        return e;
      }
      default:
        fail("an expression");
    }
  }
};

}  // namespace

std::unique_ptr<Script> parse(const std::string &src, std::string *error) {
  try {
    Parser p(lex(src));
    return p.script();
  } catch (const ParseError &e) {
    if (error) *error = std::string("Parse error: ") + e.what() + " on line " + std::to_string(e.line);
    return nullptr;
  }
}

}  // namespace ablc
