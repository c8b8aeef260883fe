// User-function analysis (a small tokenizer in place of the reference's
// libclang rewriter, libraries/stooling/src/SourceCode.cpp:98-283).
#include "fusion.h"

#include <cctype>
#include <cstdlib>
#include <map>

namespace skl {

std::string strip_comments(const std::string& s)
{
  std::string out;
  out.reserve(s.size());
  size_t i = 0, n = s.size();
  while (i < n) {
    char c = s[i];
    if (c == '"' || c == '\'') {  // literal
      char q = c;
      out += c;
      ++i;
      while (i < n && s[i] != q) {
        if (s[i] == '\\' && i + 1 < n) out += s[i++];
        out += s[i++];
      }
      if (i < n) out += s[i++];
    } else if (c == '/' && i + 1 < n && s[i + 1] == '/') {
      while (i < n && s[i] != '\n') ++i;
    } else if (c == '/' && i + 1 < n && s[i + 1] == '*') {
      i += 2;
      while (i + 1 < n && !(s[i] == '*' && s[i + 1] == '/')) {
        if (s[i] == '\n') out += '\n';
        ++i;
      }
      i += 2;
      out += ' ';
    } else {
      out += c;
      ++i;
    }
  }
  return out;
}

namespace {

struct Tok {
  enum Kind { Ident, Number, Punct, Literal, Preproc } kind;
  std::string text;
  size_t pos;  // offset in the comment-free text
};

std::vector<Tok> tokenize(const std::string& s)
{
  std::vector<Tok> toks;
  size_t i = 0, n = s.size();
  bool line_start = true;
  while (i < n) {
    char c = s[i];
    if (c == '\n') {
      line_start = true;
      ++i;
      continue;
    }
    if (isspace((unsigned char)c)) {
      ++i;
      continue;
    }
    if (c == '#' && line_start) {  // whole preprocessor line (with \-joins)
      size_t b = i;
      while (i < n && s[i] != '\n') {
        if (s[i] == '\\' && i + 1 < n && s[i + 1] == '\n') i += 2;
        else ++i;
      }
      toks.push_back({Tok::Preproc, s.substr(b, i - b), b});
      continue;
    }
    line_start = false;
    if (isalpha((unsigned char)c) || c == '_') {
      size_t b = i;
      while (i < n && (isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
      toks.push_back({Tok::Ident, s.substr(b, i - b), b});
    } else if (isdigit((unsigned char)c) ||
               (c == '.' && i + 1 < n && isdigit((unsigned char)s[i + 1]))) {
      size_t b = i;
      while (i < n && (isalnum((unsigned char)s[i]) || s[i] == '.' || s[i] == '_' ||
                       ((s[i] == '+' || s[i] == '-') &&
                        (s[i - 1] == 'e' || s[i - 1] == 'E' || s[i - 1] == 'p' ||
                         s[i - 1] == 'P'))))
        ++i;
      toks.push_back({Tok::Number, s.substr(b, i - b), b});
    } else if (c == '"' || c == '\'') {
      size_t b = i;
      char q = c;
      ++i;
      while (i < n && s[i] != q) {
        if (s[i] == '\\') ++i;
        ++i;
      }
      ++i;
      toks.push_back({Tok::Literal, s.substr(b, i - b), b});
    } else {
      static const char* two[] = {"->", "++", "--", "<<", ">>", "<=", ">=", "==",
                                  "!=", "&&", "||", "+=", "-=", "*=", "/=", "::"};
      std::string t(1, c);
      if (i + 1 < n) {
        std::string p = s.substr(i, 2);
        for (auto* x : two)
          if (p == x) t = p;
      }
      i += t.size();
      toks.push_back({Tok::Punct, t, i - t.size()});
    }
  }
  return toks;
}

std::string trim(const std::string& s)
{
  size_t b = 0, e = s.size();
  while (b < e && isspace((unsigned char)s[b])) ++b;
  while (e > b && isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

std::string collapse_ws(const std::string& s)
{
  std::string out;
  bool sp = false;
  for (char c : s) {
    if (isspace((unsigned char)c)) {
      sp = true;
    } else {
      if (sp && !out.empty()) out += ' ';
      out += c;
      sp = false;
    }
  }
  return out;
}

bool analyse_param(const std::string& decl_in, Param* p, std::string* err)
{
  p->decl = trim(decl_in);
  auto toks = tokenize(p->decl);
  int last_ident = -1;
  for (int i = (int)toks.size() - 1; i >= 0; --i) {
    if (toks[i].kind == Tok::Punct && toks[i].text == "]") {
      *err = "array parameters are not supported: '" + p->decl + "'";
      return false;
    }
    if (toks[i].kind == Tok::Ident) {
      last_ident = i;
      break;
    }
  }
  if (last_ident < 0) {
    *err = "cannot read parameter '" + p->decl + "'";
    return false;
  }
  // a lone type name (no parameter name), e.g. "float" in a prototype
  bool has_type_before = false;
  for (int i = 0; i < last_ident; ++i)
    if (toks[i].kind == Tok::Ident || toks[i].text == "*") has_type_before = true;
  if (!has_type_before) {
    *err = "parameter '" + p->decl + "' has no name";
    return false;
  }
  p->name = toks[last_ident].text;
  std::string type;
  bool is_local = false, is_ptr = false, is_matrix = false;
  for (int i = 0; i < (int)toks.size(); ++i) {
    if (i == last_ident) continue;
    const std::string& t = toks[i].text;
    if (t == "__local" || t == "local") {
      is_local = true;
      continue;  // the qualifier is not a HIP type
    }
    if (t == "*") is_ptr = true;
    if (toks[i].kind == Tok::Ident && t.size() > 9 &&
        t.compare(t.size() - 9, 9, "_matrix_t") == 0)
      is_matrix = true;
    type += (type.empty() ? "" : " ") + t;
  }
  p->type_text = type;
  if (is_local && is_ptr) {
    p->kind = ParamKind::Local;
  } else if (is_ptr) {
    p->kind = ParamKind::Buffer;
  } else if (is_matrix) {
    p->kind = ParamKind::Matrix;
  } else if (is_local) {
    *err = "__local parameter '" + p->decl + "' must be a pointer";
    return false;
  } else {
    p->kind = ParamKind::Value;
    p->value_size = known_type_size(type);
  }
  if (is_ptr) {
    std::string pt;
    for (auto& t : tokenize(type))
      if (t.text != "*" && t.text != "__global" && t.text != "global" &&
          t.text != "__constant" && t.text != "constant" && t.text != "const" &&
          t.text != "restrict" && t.text != "__restrict__")
        pt += (pt.empty() ? "" : " ") + t.text;
    p->pointee = pt;
  }
  return true;
}

}  // namespace

size_t known_type_size(const std::string& type_in)
{
  std::string t;
  for (auto& tok : tokenize(type_in))
    if (tok.text != "const" && tok.text != "__private" && tok.text != "volatile")
      t += (t.empty() ? "" : " ") + tok.text;
  static const std::map<std::string, size_t> base = {
      {"char", 1}, {"signed char", 1}, {"uchar", 1}, {"unsigned char", 1},
      {"bool", 1}, {"short", 2}, {"ushort", 2}, {"unsigned short", 2},
      {"int", 4}, {"signed int", 4}, {"uint", 4}, {"unsigned", 4},
      {"unsigned int", 4}, {"float", 4}, {"long", 8}, {"ulong", 8},
      {"unsigned long", 8}, {"long long", 8}, {"unsigned long long", 8},
      {"double", 8}, {"size_t", 8}, {"Index", 8}, {"IndexPoint", 16}};
  auto it = base.find(t);
  if (it != base.end()) return it->second;
  static const char* vec_bases[] = {"char", "uchar", "short", "ushort", "int",
                                    "uint", "long", "ulong", "float", "double"};
  for (auto* b : vec_bases) {
    std::string bs(b);
    if (t.size() > bs.size() && t.compare(0, bs.size(), bs) == 0) {
      std::string rest = t.substr(bs.size());
      int nn = 0;
      if (rest == "2") nn = 2;
      else if (rest == "3" || rest == "4") nn = 4;
      else if (rest == "8") nn = 8;
      else if (rest == "16") nn = 16;
      if (nn) return base.at(bs) * (size_t)nn;
    }
  }
  return 0;
}

bool parse_user_function(const std::string& src, const std::string& name,
                         UserFunction* out, std::string* err)
{
  const std::string s = strip_comments(src);
  auto toks = tokenize(s);
  for (size_t i = 0; i + 1 < toks.size(); ++i) {
    if (toks[i].kind != Tok::Ident || toks[i].text != name) continue;
    if (toks[i + 1].text != "(") continue;
    if (i > 0 && (toks[i - 1].text == "." || toks[i - 1].text == "->")) continue;
    // matching ')'
    size_t j = i + 1;
    int depth = 0;
    for (; j < toks.size(); ++j) {
      if (toks[j].text == "(") ++depth;
      else if (toks[j].text == ")" && --depth == 0) break;
    }
    if (j >= toks.size()) break;
    if (j + 1 >= toks.size() || toks[j + 1].text != "{") continue;  // not a def
    // body
    size_t k = j + 1;
    depth = 0;
    for (; k < toks.size(); ++k) {
      if (toks[k].text == "{") ++depth;
      else if (toks[k].text == "}" && --depth == 0) break;
    }
    if (k >= toks.size()) {
      *err = "unbalanced braces in the definition of '" + name + "'";
      return false;
    }
    // return type: back to the previous statement boundary
    size_t b = i;
    while (b > 0) {
      const Tok& t = toks[b - 1];
      if (t.kind == Tok::Preproc || t.text == ";" || t.text == "}" || t.text == "{")
        break;
      --b;
    }
    std::string ret;
    for (size_t x = b; x < i; ++x) {
      if (toks[x].text == "inline" || toks[x].text == "static" ||
          toks[x].text == "__kernel" || toks[x].text == "__device__" ||
          toks[x].text == "__forceinline__")
        continue;
      ret += (ret.empty() ? "" : " ") + toks[x].text;
    }
    if (ret.empty()) continue;  // a call, not a definition
    UserFunction f;
    f.name = name;
    f.return_type = ret;
    // parameters: split at depth-0 commas
    std::string plist = s.substr(toks[i + 1].pos + 1, toks[j].pos - toks[i + 1].pos - 1);
    std::vector<std::string> parts;
    std::string cur;
    depth = 0;
    for (char c : plist) {
      if (c == '(' || c == '[' || c == '<') ++depth;
      if (c == ')' || c == ']' || c == '>') --depth;
      if (c == ',' && depth == 0) {
        parts.push_back(cur);
        cur.clear();
      } else {
        cur += c;
      }
    }
    if (!trim(cur).empty() || !parts.empty()) parts.push_back(cur);
    if (parts.size() == 1 && trim(parts[0]) == "void") parts.clear();
    for (auto& part : parts) {
      Param p;
      if (!analyse_param(part, &p, err)) {
        *err = "in '" + name + "': " + *err;
        return false;
      }
      f.params.push_back(p);
    }
    f.body = s.substr(toks[j + 1].pos + 1, toks[k].pos - toks[j + 1].pos - 1);
    *out = f;
    return true;
  }
  *err = "could not find a definition of the user function '" + name + "'";
  return false;
}

UserOp classify_binary(const UserFunction& f)
{
  if (f.params.size() != 2) return UserOp::Other;
  std::string body;
  for (char c : f.body)
    if (!isspace((unsigned char)c)) body += c;
  const std::string& a = f.params[0].name;
  const std::string& b = f.params[1].name;
  auto is = [&](const std::string& expr) {
    return body == "return" + expr + ";" || body == "return(" + expr + ");";
  };
  if (is(a + "+" + b) || is(b + "+" + a)) return UserOp::Add;
  if (is(a + "*" + b) || is(b + "*" + a)) return UserOp::Mul;
  if (is("max(" + a + "," + b + ")") || is("fmax(" + a + "," + b + ")") ||
      is("fmaxf(" + a + "," + b + ")") || is(a + ">" + b + "?" + a + ":" + b))
    return UserOp::Max;
  if (is("min(" + a + "," + b + ")") || is("fmin(" + a + "," + b + ")") ||
      is("fminf(" + a + "," + b + ")") || is(a + "<" + b + "?" + a + ":" + b))
    return UserOp::Min;
  return UserOp::Other;
}

int match_weighted_stencil(const std::string& src, const UserFunction& f)
{
  for (auto& t : tokenize(strip_comments(src)))
    if (t.kind == Tok::Preproc) return -1;
  if (collapse_ws(f.return_type) != "float" || f.params.size() != 2) return -1;
  const Param& pi = f.params[0];
  const Param& pw = f.params[1];
  if (collapse_ws(pi.type_text) != "input_matrix_t" || pi.name.empty()) return -1;
  if (pw.kind != ParamKind::Buffer || collapse_ws(pw.pointee) != "float" || pw.name.empty())
    return -1;
  for (auto& t : tokenize(pw.type_text))
    if (t.text != "__global" && t.text != "global" && t.text != "const" && t.text != "float" &&
        t.text != "*")
      return -1;
  const std::vector<Tok> tk = tokenize(f.body);
  size_t at = 0;
  auto lit = [&](const char* x) {
    if (at < tk.size() && tk[at].text == x) {
      ++at;
      return true;
    }
    return false;
  };
  auto seq = [&](std::initializer_list<const char*> xs) {
    const size_t save = at;
    for (const char* x : xs)
      if (!lit(x)) {
        at = save;
        return false;
      }
    return true;
  };
  auto ident = [&](std::string* out) {
    if (at < tk.size() && tk[at].kind == Tok::Ident) {
      *out = tk[at++].text;
      return true;
    }
    return false;
  };
  auto number = [&](long* out) {
    if (at >= tk.size() || tk[at].kind != Tok::Number) return false;
    const std::string& x = tk[at].text;
    for (char c : x)
      if (!isdigit((unsigned char)c)) return false;
    *out = std::strtol(x.c_str(), nullptr, 10);
    ++at;
    return true;
  };
  auto zero = [&]() {
    if (at >= tk.size() || tk[at].kind != Tok::Number) return false;
    std::string x = tk[at].text;
    if (!x.empty() && (x.back() == 'f' || x.back() == 'F')) x.pop_back();
    if (x.empty() || x.find_first_not_of("0.") != std::string::npos ||
        x.find_first_of("0") == std::string::npos)
      return false;
    ++at;
    return true;
  };
  std::string S, Y, X, I, Wn;
  long n = -1, m = 0;
  // float S = 0.0f ;
  if (!lit("float") || !ident(&S) || !lit("=") || !zero() || !lit(";")) return -1;
  // for ( int V = - N ; V <= N ; ++V | V++ )
  auto loop = [&](std::string* v) {
    long a, b;
    std::string v2, v3;
    if (!seq({"for", "(", "int"}) || !ident(v) || !lit("=") || !lit("-") || !number(&a) ||
        !lit(";") || !ident(&v2) || v2 != *v || !lit("<=") || !number(&b) || !lit(";"))
      return false;
    if (a != b || (n >= 0 && a != n)) return false;
    n = a;
    if (lit("++")) {
      if (!ident(&v3) || v3 != *v) return false;
    } else if (!ident(&v3) || v3 != *v || !lit("++")) {
      return false;
    }
    return lit(")");
  };
  if (!loop(&Y)) return -1;
  const bool b1 = lit("{");
  if (!loop(&X)) return -1;
  const bool b2 = lit("{");
  // S += P ;  or  S = S + P ;
  std::string s2;
  if (!ident(&s2) || s2 != S) return -1;
  if (!lit("+=")) {
    std::string s3;
    if (!lit("=") || !ident(&s3) || s3 != S || !lit("+")) return -1;
  }
  auto get = [&]() {
    std::string g, i, x, y;
    if (!ident(&g) || g != "getData" || !lit("(") || !ident(&i) || !lit(",") || !ident(&x) ||
        !lit(",") || !ident(&y) || !lit(")"))
      return false;
    return i == pi.name && x == X && y == Y;
  };
  auto plus_n = [&](const std::string& v, bool paren) {
    std::string u;
    long k;
    if (paren && !lit("(")) return false;
    if (!ident(&u) || u != v || !lit("+") || !number(&k) || k != n) return false;
    return !paren || lit(")");
  };
  auto weight = [&]() {
    std::string wn;
    if (!ident(&wn) || wn != pw.name || !lit("[")) return false;
    const size_t save = at;
    bool ok = plus_n(Y, true) && lit("*") && number(&m);
    if (!ok) {
      at = save;
      ok = number(&m) && lit("*") && plus_n(Y, true);
    }
    if (!ok || m != 2 * n + 1 || !lit("+")) return false;
    const size_t save2 = at;
    if (!plus_n(X, true)) {
      at = save2;
      if (!plus_n(X, false)) return false;
    }
    return lit("]");
  };
  const size_t pstart = at;
  if (!(get() && lit("*") && weight())) {
    at = pstart;
    if (!(weight() && lit("*") && get())) return -1;
  }
  if (!lit(";")) return -1;
  if (b2 && !lit("}")) return -1;
  if (b1 && !lit("}")) return -1;
  std::string s4;
  if (!lit("return") || !ident(&s4) || s4 != S || !lit(";") || at != tk.size()) return -1;
  // distinct names, none a keyword the pattern relies on
  const std::string names[] = {S, X, Y, pi.name, pw.name};
  for (int a = 0; a < 5; ++a) {
    if (names[a] == "getData" || names[a] == "float" || names[a] == "int") return -1;
    for (int b = a + 1; b < 5; ++b)
      if (names[a] == names[b]) return -1;
  }
  return (int)n;
}

bool uses_work_item_builtins(const std::string& src)
{
  static const char* names[] = {"get_global_id", "get_local_id", "get_group_id",
                                "get_local_size", "get_num_groups",
                                "get_global_size", "barrier", "mem_fence",
                                "__local", "local"};
  for (auto& t : tokenize(strip_comments(src)))
    if (t.kind == Tok::Ident)
      for (auto* nm : names)
        if (t.text == nm) return true;
  return false;
}

ExtraPlumbing make_extra_plumbing(const std::vector<Param>& extras,
                                  const std::string& prefix)
{
  ExtraPlumbing e;
  for (size_t i = 0; i < extras.size(); ++i) {
    const Param& p = extras[i];
    std::string nm = "skl_" + prefix + "xa" + std::to_string(i);
    if (p.kind == ParamKind::Local) {
      e.params += ", unsigned int " + nm + "_ofs";
      e.setup += p.type_text + " " + nm + " = (" + p.type_text + ")(skl_dyn_lds + " +
                 nm + "_ofs); ";
    } else {
      e.params += ", " + p.type_text + " " + nm;
    }
    e.args += ", " + nm;
  }
  return e;
}

}  // namespace skl
