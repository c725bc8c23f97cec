// parser.hpp -- recursive-descent parser of the .abl language.
#pragma once

#include <memory>
#include <string>

#include "ast.hpp"

namespace ablc {

// Parses `src`; on a syntax error returns nullptr and sets *error to
// "Parse error: <message> on line <n>".
std::unique_ptr<Script> parse(const std::string &src, std::string *error);

}  // namespace ablc
