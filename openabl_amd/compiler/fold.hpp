// fold.hpp -- folding of compile-time values (fold.cpp).
#pragma once

#include <string>
#include <vector>

#include "ast.hpp"

namespace ablc {

Value fold_unary(UnOp op, const Value &v);
Value fold_binary(BinOp op, const Value &l, const Value &r);
// one-argument math builtins (sin ... round) and pow, by signature name
Value fold_builtin(const std::string &sig_name, const std::vector<Value> &args);
// literal expression of a folded value (nullptr if the value is invalid)
ExprP value_expr(const Value &v);

}  // namespace ablc
