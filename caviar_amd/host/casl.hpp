// CASL reader for the objects of the LJ + Ewald path.
//
// The token rules follow the reference interpreter (interpreter/input/lexer.cpp,
// parser.cpp): '#' starts a comment, a line end is a token, numbers are int
// unless they carry '.' or an exponent, a leading '+'/'-' folds into the
// number that follows (Parser::get_val_token, parser.cpp:145-185), strings are
// quoted with either quote character, identifiers may contain '.' (domain's "lower_global.x").
// Values that accept a variable resolve `int`/`real` names the way the
// reference's GET_OR_CHOOSE_A_REAL / GET_OR_CHOOSE_A_INT macros do
// (interpreter/object_handler/preprocessors.h:192-300).
#pragma once

#include <map>
#include <stdexcept>
#include <string>

namespace caviar_amd {

// Raised for every script or setup error; the driver prints it and exits with
// status 1, as the reference's error->all does (interpreter/error.cpp:175-202).
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

namespace casl {

enum class Kind { identifier, string, int_number, real_number, plus, minus, mul, div, lparen, rparen,
                  assign, eol, eof, other };

struct Token {
  Kind kind = Kind::eof;
  std::string text;  // identifier / string contents / the raw characters
  long ivalue = 0;
  double rvalue = 0.0;
  int line = 0;
};

// `int NAME value` and `real NAME value` definitions (object_creator/basic_types.cpp).
struct Variables {
  std::map<std::string, long> ints;
  std::map<std::string, double> reals;
};

class Parser {
 public:
  // src: the script text; name: file name for messages; dir: the directory that
  // relative file names in the script are resolved against.
  Parser(std::string src, std::string name, std::string dir, Variables *vars);
  static Parser from_file(const std::string &path, Variables *vars);

  Token raw();           // next token as written
  Token value();         // numbers with their sign folded in
  void unget(const Token &t);
  void end_of_line();    // skip the rest of the current line

  // A real: literal or the name of an int/real variable.
  double real(const std::string &what);
  // An int: literal or the name of an int variable.
  long integer(const std::string &what);
  // An object name (identifier).
  std::string name(const std::string &what);
  // A quoted string.
  std::string string_literal(const std::string &what);
  // + - * / ( ) over numbers and variables, up to the end of the line.
  double expression();

  [[noreturn]] void fail(const std::string &msg) const;
  const std::string &dir() const { return dir_; }
  Variables *vars() const { return vars_; }

 private:
  Token lex();
  double expr_sum();
  double expr_product();
  double expr_primary();

  std::string s_, name_, dir_;
  size_t p_ = 0;
  int line_ = 1;
  Variables *vars_;
  bool has_back_ = false;
  Token back_;
};

}  // namespace casl
}  // namespace caviar_amd
