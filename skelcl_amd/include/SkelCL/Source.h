// skelcl::Source (Source.h, src/Source.cpp:51-100): user source text from a
// string or a stream; an empty source is an error.  detail::CommonDefinitions
// (Source.h:122-166, src/Source.cpp:102-142): text that every skeleton kernel
// built afterwards starts with; the library keeps it (one registry per
// process, see skl_common_definition_append).
#pragma once

#include <istream>
#include <iterator>
#include <sstream>
#include <string>

#include "detail/Runtime.h"

namespace skelcl {

class Source {
public:
  Source() = default;
  Source(const char* s) : _s(s ? s : "") { checkNotEmpty(); }
  Source(const std::string& s) : _s(s) { checkNotEmpty(); }
  Source(std::istream& is)
      : _s(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>())
  {
    checkNotEmpty();
  }
  Source(std::istream&& is)
      : _s(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>())
  {
    checkNotEmpty();
  }
  operator std::string() const { return _s; }
  const std::string& str() const { return _s; }
  void append(const std::string& more) { _s.append("\n" + more); }

private:
  void checkNotEmpty() const
  {
    if (_s.empty())
      detail::abortWithError("Tried to create source object with empty user source.");
  }
  std::string _s;
};

namespace detail {

class CommonDefinitions {
public:
  enum Level : unsigned int { PRAGMA, USER_DEFINITION, GENERATED_DEFINITION, SIZE };

  static void append(const std::string& source, Level level)
  {
    check(skl_common_definition_append(source.c_str(), (int)level),
          "CommonDefinitions::append");
  }
  // every level's text, in level order
  static Source getSource()
  {
    char* t = nullptr;
    check(skl_common_definitions(&t), "CommonDefinitions::getSource");
    Source s;
    if (t && *t) s.append(t);
    skl_free_string(t);
    return s;
  }

private:
  CommonDefinitions() = delete;
};

class RegisterCommonDefinition {
public:
  RegisterCommonDefinition(const char* definition,
                           CommonDefinitions::Level level = CommonDefinitions::USER_DEFINITION)
  {
    CommonDefinitions::append(definition, level);
  }
};

class RegisterCommonMacroDefinition {
public:
  template <typename T>
  RegisterCommonMacroDefinition(const char* name, T&& value)
  {
    std::ostringstream o;
    o << "#define " << name << " " << value;
    CommonDefinitions::append(o.str(), CommonDefinitions::USER_DEFINITION);
  }
};

}  // namespace detail
}  // namespace skelcl
