#include "casl.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace caviar_amd {
namespace casl {

Parser::Parser(std::string src, std::string name, std::string dir, Variables *vars)
    : s_(std::move(src)), name_(std::move(name)), dir_(std::move(dir)), vars_(vars)
{
}

Parser Parser::from_file(const std::string &path, Variables *vars)
{
  std::ifstream f(path);
  if (!f) throw Error("file does not exist : " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const size_t slash = path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
  return Parser(ss.str(), path, dir, vars);
}

void Parser::fail(const std::string &msg) const
{
  throw Error(name_ + ":" + std::to_string(line_) + ": " + msg);
}

static bool ident_start(char c) { return std::isalpha((unsigned char)c) || c == '_'; }
static bool ident_char(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '.'; }

Token Parser::lex()
{
  Token t;
  for (;;) {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\r')) ++p_;
    if (p_ < s_.size() && s_[p_] == '#') {
      while (p_ < s_.size() && s_[p_] != '\n') ++p_;
    }
    break;
  }
  t.line = line_;
  if (p_ >= s_.size()) {
    t.kind = Kind::eof;
    return t;
  }
  const char c = s_[p_];
  if (c == '\n') {
    ++p_;
    ++line_;
    t.kind = Kind::eol;
    return t;
  }
  if (ident_start(c)) {
    const size_t b = p_;
    while (p_ < s_.size() && ident_char(s_[p_])) ++p_;
    t.kind = Kind::identifier;
    t.text = s_.substr(b, p_ - b);
    return t;
  }
  if (std::isdigit((unsigned char)c) ||
      (c == '.' && p_ + 1 < s_.size() && std::isdigit((unsigned char)s_[p_ + 1]))) {
    const size_t b = p_;
    bool real = false;
    while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
    if (p_ < s_.size() && s_[p_] == '.') {
      real = true;
      ++p_;
      while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
    }
    if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
      size_t q = p_ + 1;
      if (q < s_.size() && (s_[q] == '+' || s_[q] == '-')) ++q;
      if (q < s_.size() && std::isdigit((unsigned char)s_[q])) {
        real = true;
        p_ = q;
        while (p_ < s_.size() && std::isdigit((unsigned char)s_[p_])) ++p_;
      }
    }
    t.text = s_.substr(b, p_ - b);
    if (real) {
      t.kind = Kind::real_number;
      t.rvalue = std::strtod(t.text.c_str(), nullptr);
    } else {
      t.kind = Kind::int_number;
      t.ivalue = std::strtol(t.text.c_str(), nullptr, 10);
    }
    return t;
  }
  if (c == '"' || c == '\'') {  // either quote opens a string (lexer.cpp:319-321)
    const size_t b = ++p_;
    while (p_ < s_.size() && s_[p_] != c && s_[p_] != '\n') ++p_;
    if (p_ >= s_.size() || s_[p_] != c) fail("unterminated string");
    t.kind = Kind::string;
    t.text = s_.substr(b, p_ - b);
    ++p_;
    return t;
  }
  ++p_;
  t.text = std::string(1, c);
  switch (c) {
    case '+': t.kind = Kind::plus; break;
    case '-': t.kind = Kind::minus; break;
    case '*': t.kind = Kind::mul; break;
    case '/': t.kind = Kind::div; break;
    case '(': t.kind = Kind::lparen; break;
    case ')': t.kind = Kind::rparen; break;
    case '=': t.kind = Kind::assign; break;
    default: t.kind = Kind::other; break;
  }
  return t;
}

Token Parser::raw()
{
  if (has_back_) {
    has_back_ = false;
    return back_;
  }
  return lex();
}

void Parser::unget(const Token &t)
{
  back_ = t;
  has_back_ = true;
}

Token Parser::value()
{
  Token t = raw();
  if (t.kind != Kind::plus && t.kind != Kind::minus) return t;
  const bool neg = t.kind == Kind::minus;
  Token n = raw();
  if (n.kind == Kind::int_number) {
    if (neg) n.ivalue = -n.ivalue;
  } else if (n.kind == Kind::real_number) {
    if (neg) n.rvalue = -n.rvalue;
  } else {
    fail("Unexpected token");
  }
  return n;
}

void Parser::end_of_line()
{
  for (;;) {
    Token t = raw();
    if (t.kind == Kind::eol || t.kind == Kind::eof) return;
  }
}

double Parser::real(const std::string &what)
{
  Token t = value();
  if (t.kind == Kind::real_number) return t.rvalue;
  if (t.kind == Kind::int_number) return (double)t.ivalue;
  if (t.kind == Kind::identifier) {
    auto i = vars_->ints.find(t.text);
    if (i != vars_->ints.end()) return (double)i->second;
    auto r = vars_->reals.find(t.text);
    if (r != vars_->reals.end()) return r->second;
    fail("Invalid object NAME " + what + ": " + t.text);
  }
  fail("Expected a real number as " + what);
}

long Parser::integer(const std::string &what)
{
  Token t = value();
  if (t.kind == Kind::int_number) return t.ivalue;
  if (t.kind == Kind::identifier) {
    auto i = vars_->ints.find(t.text);
    if (i != vars_->ints.end()) return i->second;
    fail(t.text + " is not referred to an int number (" + what + ")");
  }
  fail("Expected an int number as " + what);
}

std::string Parser::name(const std::string &what)
{
  Token t = value();
  if (t.kind != Kind::identifier) fail("Expected an object NAME as " + what);
  return t.text;
}

std::string Parser::string_literal(const std::string &what)
{
  Token t = value();
  if (t.kind != Kind::string) fail("Expected a string as " + what);
  return t.text;
}

double Parser::expression() { return expr_sum(); }

double Parser::expr_sum()
{
  double v = expr_product();
  for (;;) {
    Token t = raw();
    if (t.kind == Kind::plus)
      v += expr_product();
    else if (t.kind == Kind::minus)
      v -= expr_product();
    else {
      unget(t);
      return v;
    }
  }
}

double Parser::expr_product()
{
  double v = expr_primary();
  for (;;) {
    Token t = raw();
    if (t.kind == Kind::mul)
      v *= expr_primary();
    else if (t.kind == Kind::div) {
      const double d = expr_primary();
      if (d == 0.0) fail("divide by zero");
      v /= d;
    } else {
      unget(t);
      return v;
    }
  }
}

double Parser::expr_primary()
{
  Token t = raw();
  switch (t.kind) {
    case Kind::int_number: return (double)t.ivalue;
    case Kind::real_number: return t.rvalue;
    case Kind::minus: return -expr_primary();
    case Kind::plus: return expr_primary();
    case Kind::lparen: {
      const double v = expr_sum();
      if (raw().kind != Kind::rparen) fail("')' expected");
      return v;
    }
    case Kind::identifier: {
      auto i = vars_->ints.find(t.text);
      if (i != vars_->ints.end()) return (double)i->second;
      auto r = vars_->reals.find(t.text);
      if (r != vars_->reals.end()) return r->second;
      fail("undefined variable: " + t.text);
    }
    default: fail("primary expected");
  }
}

}  // namespace casl
}  // namespace caviar_amd
