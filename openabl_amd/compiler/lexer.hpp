// lexer.hpp -- tokens of the .abl language.
#pragma once

#include <string>
#include <vector>

namespace ablc {

enum class Tok {
  End, Agent, Break, Continue, Else, Environment, If, For, New, Param, Position, Return,
  Sequential, Simulate, Step, While,
  Add, Sub, Mul, Div, Mod, BitAnd, BitXor, BitOr, Shl, Shr, Not, BitNot, Qm, Dot, DotDot, Comma,
  Colon, Semi, LParen, RParen, LBracket, RBracket, LBrace, RBrace, Assign, Arrow, Eq, Ne, Le, Ge,
  Lt, Gt, And, Or, AddAssign, SubAssign, MulAssign, DivAssign, ModAssign, AndAssign, XorAssign,
  OrAssign, ShlAssign, ShrAssign,
  Bool, Int, Float, String, Ident,
  Error,
};

struct Token {
  Tok kind = Tok::End;
  std::string text;  // identifier / string contents / the token's source text
  long ival = 0;
  double fval = 0.0;
  bool bval = false;
  int line = 1;      // line the token begins on (see lexer.cpp)
  int end_line = 1;  // line it ends on
};

// Whole-file scan; a character no rule matches yields a Tok::Error token.
std::vector<Token> lex(const std::string &src);

const char *token_name(Tok t);

}  // namespace ablc
