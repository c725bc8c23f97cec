// lexer.cpp -- scanner of the .abl language.
//
// Tokens and literal rules follow the reference scanner (src/Lexer.l:62-161):
// keywords, `true`/`false`, decimal and hex integers, floats with a dot
// (`1.`, `.5`, `1.5e-3`; `1e5` is an integer followed by an identifier), an
// integer directly followed by `..` is an integer (so `0..n` is a range),
// strings whose backslash escapes drop the backslash, `//` and `/* */`
// comments.
//
// Line numbers reproduce where that scanner's locations begin, since the
// diagnostics the analysis prints carry them: a token begins where the
// previous token ended, unless a run of blanks (spaces, tabs, carriage
// returns) or the end of a block comment came in between, which moves the
// beginning to there.  Newlines and line comments do not move it, so a
// token at the start of a line that follows a blank line without
// indentation "begins" on the line where the previous token ended.
#include "lexer.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>

namespace ablc {

namespace {

struct Keyword {
  const char *text;
  Tok tok;
};

const Keyword kKeywords[] = {
    {"agent", Tok::Agent},       {"break", Tok::Break},         {"continue", Tok::Continue},
    {"else", Tok::Else},         {"environment", Tok::Environment}, {"if", Tok::If},
    {"for", Tok::For},           {"new", Tok::New},             {"param", Tok::Param},
    {"position", Tok::Position}, {"return", Tok::Return},       {"sequential", Tok::Sequential},
    {"simulate", Tok::Simulate}, {"step", Tok::Step},           {"while", Tok::While},
};

// Operators, longest first so that a prefix never wins.
const Keyword kOps[] = {
    {"<<=", Tok::ShlAssign}, {">>=", Tok::ShrAssign}, {"..", Tok::DotDot}, {"->", Tok::Arrow},
    {"==", Tok::Eq},         {"!=", Tok::Ne},         {"<=", Tok::Le},     {">=", Tok::Ge},
    {"<<", Tok::Shl},        {">>", Tok::Shr},        {"&&", Tok::And},    {"||", Tok::Or},
    {"+=", Tok::AddAssign},  {"-=", Tok::SubAssign},  {"*=", Tok::MulAssign},
    {"/=", Tok::DivAssign},  {"%=", Tok::ModAssign},  {"&=", Tok::AndAssign},
    {"^=", Tok::XorAssign},  {"|=", Tok::OrAssign},   {"+", Tok::Add},     {"-", Tok::Sub},
    {"*", Tok::Mul},         {"/", Tok::Div},         {"%", Tok::Mod},     {"&", Tok::BitAnd},
    {"^", Tok::BitXor},      {"|", Tok::BitOr},       {"=", Tok::Assign},  {"!", Tok::Not},
    {"~", Tok::BitNot},      {"?", Tok::Qm},          {".", Tok::Dot},     {",", Tok::Comma},
    {":", Tok::Colon},       {";", Tok::Semi},        {"<", Tok::Lt},      {">", Tok::Gt},
    {"(", Tok::LParen},      {")", Tok::RParen},      {"[", Tok::LBracket}, {"]", Tok::RBracket},
    {"{", Tok::LBrace},      {"}", Tok::RBrace},
};

bool is_id_start(char c) { return std::isalpha((unsigned char) c) || c == '_'; }
bool is_id_char(char c) { return std::isalnum((unsigned char) c) || c == '_'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

}  // namespace

std::vector<Token> lex(const std::string &src) {
  std::vector<Token> out;
  size_t p = 0;
  const size_t n = src.size();
  int line = 1;        // current (end) line
  int begin_line = 1;  // where the next token begins
  auto emit = [&](Token t, size_t len) {
    // newlines inside the token (strings) advance the end line
    for (size_t k = 0; k < len; k++)
      if (src[p + k] == '\n') line++;
    t.line = begin_line;
    t.end_line = line;
    p += len;
    begin_line = line;  // the next scan starts where this token ended
    out.push_back(std::move(t));
  };
  while (p < n) {
    const char c = src[p];
    if (c == ' ' || c == '\t' || c == '\r') {
      while (p < n && (src[p] == ' ' || src[p] == '\t' || src[p] == '\r')) p++;
      begin_line = line;
      continue;
    }
    if (c == '\n') {
      while (p < n && src[p] == '\n') {
        p++;
        line++;
      }
      continue;
    }
    if (c == '/' && p + 1 < n && src[p + 1] == '/') {
      while (p < n && src[p] != '\n') p++;
      continue;
    }
    if (c == '/' && p + 1 < n && src[p + 1] == '*') {
      size_t q = p + 2;
      while (q < n && !(src[q] == '*' && q + 1 < n && src[q + 1] == '/')) {
        if (src[q] == '\n') line++;
        q++;
      }
      p = q + 2 <= n ? q + 2 : n;
      begin_line = line;
      continue;
    }
    Token t;
    if (is_id_start(c)) {
      size_t q = p;
      while (q < n && is_id_char(src[q])) q++;
      const std::string word = src.substr(p, q - p);
      t.kind = Tok::Ident;
      t.text = word;
      for (const Keyword &k : kKeywords)
        if (word == k.text) t.kind = k.tok;
      if (word == "true" || word == "false") {
        t.kind = Tok::Bool;
        t.bval = word == "true";
      }
      emit(t, q - p);
      continue;
    }
    if (is_digit(c) || (c == '.' && p + 1 < n && is_digit(src[p + 1]))) {
      // hex integer
      if (c == '0' && p + 1 < n && (src[p + 1] == 'x' || src[p + 1] == 'X') && p + 2 < n &&
          std::isxdigit((unsigned char) src[p + 2])) {
        size_t q = p + 2;
        while (q < n && std::isxdigit((unsigned char) src[q])) q++;
        t.kind = Tok::Int;
        t.text = src.substr(p, q - p);
        t.ival = std::strtol(t.text.c_str(), nullptr, 16);
        emit(t, q - p);
        continue;
      }
      size_t q = p;
      while (q < n && is_digit(src[q])) q++;
      const bool had_int = q > p;
      // an integer directly followed by `..` is an integer (a range start)
      if (had_int && q + 1 < n && src[q] == '.' && src[q + 1] == '.') {
        t.kind = Tok::Int;
        t.text = src.substr(p, q - p);
        t.ival = std::strtol(t.text.c_str(), nullptr, 10);
        emit(t, q - p);
        continue;
      }
      if (q < n && src[q] == '.') {
        size_t r = q + 1;
        while (r < n && is_digit(src[r])) r++;
        if (r < n && (src[r] == 'e' || src[r] == 'E')) {
          size_t e = r + 1;
          if (e < n && (src[e] == '+' || src[e] == '-')) e++;
          if (e < n && is_digit(src[e])) {
            while (e < n && is_digit(src[e])) e++;
            r = e;
          }
        }
        t.kind = Tok::Float;
        t.text = src.substr(p, r - p);
        t.fval = std::strtod(t.text.c_str(), nullptr);
        emit(t, r - p);
        continue;
      }
      t.kind = Tok::Int;
      t.text = src.substr(p, q - p);
      t.ival = std::strtol(t.text.c_str(), nullptr, 10);
      emit(t, q - p);
      continue;
    }
    if (c == '"') {
      size_t q = p + 1;
      std::string s;
      bool closed = false;
      while (q < n) {
        if (src[q] == '\\' && q + 1 < n) {
          s.push_back(src[q + 1]);
          q += 2;
          continue;
        }
        if (src[q] == '"') {
          closed = true;
          q++;
          break;
        }
        s.push_back(src[q]);
        q++;
      }
      if (!closed) {
        t.kind = Tok::Error;
        t.text = "\"";
        emit(t, 1);
        continue;
      }
      t.kind = Tok::String;
      t.text = s;
      emit(t, q - p);
      continue;
    }
    bool matched = false;
    for (const Keyword &k : kOps) {
      const size_t len = std::strlen(k.text);
      if (src.compare(p, len, k.text) == 0) {
        t.kind = k.tok;
        t.text = k.text;
        emit(t, len);
        matched = true;
        break;
      }
    }
    if (!matched) {
      t.kind = Tok::Error;
      t.text = std::string(1, c);
      emit(t, 1);
    }
  }
  Token end;
  end.kind = Tok::End;
  end.line = begin_line;
  end.end_line = line;
  out.push_back(end);
  return out;
}

const char *token_name(Tok t) {
  switch (t) {
    case Tok::End: return "end of file";
    case Tok::Ident: return "IDENTIFIER";
    case Tok::Int: return "INT";
    case Tok::Float: return "FLOAT";
    case Tok::Bool: return "BOOL";
    case Tok::String: return "STRING";
    case Tok::Error: return "invalid character";
    default: break;
  }
  for (const Keyword &k : kKeywords)
    if (k.tok == t) return k.text;
  for (const Keyword &k : kOps)
    if (k.tok == t) return k.text;
  return "token";
}

}  // namespace ablc
