// config.cpp -- readConfig (libconfig-subset reader) and AlgoParser.
// Reference: src/independant/readConfig.cc:29-163, include/readConfig.hpp:13-62.
#include "../../include/readConfig.hpp"
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace {

// ---------------------------------------------------------------------------
// libconfig subset: groups {..}, scalar settings "name = value;" or
// "name: value;", strings (adjacent literals concatenate), integers
// (decimal/hex, optional L suffix), floats, booleans (true/false in any
// case), comments //, #, /* */.  Lists/arrays are parsed and ignored.
// ---------------------------------------------------------------------------
struct Setting {
  enum Kind { INT, FLOAT, BOOL, STRING, GROUP, LIST } kind = GROUP;
  long long i = 0;
  double f = 0;
  bool b = false;
  std::string s;
  std::map<std::string, std::unique_ptr<Setting>> children;
};

struct Parser {
  const std::string& src;
  size_t pos = 0;
  int line = 1;
  std::string err;
  explicit Parser(const std::string& s) : src(s) {}

  void skip() {
    for (;;) {
      while (pos < src.size() && isspace((unsigned char)src[pos])) {
        if (src[pos] == '\n') line++;
        pos++;
      }
      if (pos + 1 < src.size() && src[pos] == '/' && src[pos + 1] == '/') {
        while (pos < src.size() && src[pos] != '\n') pos++;
      } else if (pos < src.size() && src[pos] == '#') {
        while (pos < src.size() && src[pos] != '\n') pos++;
      } else if (pos + 1 < src.size() && src[pos] == '/' && src[pos + 1] == '*') {
        size_t e = src.find("*/", pos + 2);
        for (size_t k = pos; k < (e == std::string::npos ? src.size() : e); k++)
          if (src[k] == '\n') line++;
        pos = e == std::string::npos ? src.size() : e + 2;
      } else {
        return;
      }
    }
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
    return false;
  }
  bool name(std::string& out) {
    skip();
    size_t s = pos;
    if (pos < src.size() && (isalpha((unsigned char)src[pos]) || src[pos] == '*')) {
      pos++;
      while (pos < src.size() && (isalnum((unsigned char)src[pos]) || src[pos] == '_' || src[pos] == '-' || src[pos] == '*'))
        pos++;
    }
    if (pos == s) return fail("expected a setting name");
    out = src.substr(s, pos - s);
    return true;
  }
  bool value(Setting& v) {
    skip();
    if (pos >= src.size()) return fail("unexpected end of file");
    char ch = src[pos];
    if (ch == '{') {
      pos++;
      v.kind = Setting::GROUP;
      return group_body(v, '}');
    }
    if (ch == '(' || ch == '[') {
      char close = ch == '(' ? ')' : ']';
      pos++;
      v.kind = Setting::LIST;
      for (;;) {
        skip();
        if (pos < src.size() && src[pos] == close) {
          pos++;
          return true;
        }
        Setting item;
        if (!value(item)) return false;
        skip();
        if (pos < src.size() && src[pos] == ',') pos++;
      }
    }
    if (ch == '"') {
      v.kind = Setting::STRING;
      v.s.clear();
      for (;;) {
        pos++;
        while (pos < src.size() && src[pos] != '"') {
          if (src[pos] == '\\' && pos + 1 < src.size()) {
            pos++;
            char e = src[pos];
            v.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
          } else {
            v.s += src[pos];
          }
          pos++;
        }
        if (pos >= src.size()) return fail("unterminated string");
        pos++;
        skip();
        if (pos < src.size() && src[pos] == '"') continue;  // adjacent literals
        return true;
      }
    }
    // boolean?
    size_t s = pos;
    while (pos < src.size() && (isalnum((unsigned char)src[pos]) || strchr("+-._", src[pos]))) pos++;
    std::string tok = src.substr(s, pos - s);
    if (tok.empty()) return fail(std::string("unexpected character '") + ch + "'");
    std::string low;
    for (char c : tok) low += (char)tolower((unsigned char)c);
    if (low == "true" || low == "false") {
      v.kind = Setting::BOOL;
      v.b = low == "true";
      return true;
    }
    const char* p = tok.c_str();
    char* end = nullptr;
    bool isfloat = tok.find_first_of(".eE") != std::string::npos && !(low.size() > 1 && low[0] == '0' && low[1] == 'x');
    if (isfloat) {
      v.kind = Setting::FLOAT;
      v.f = strtod(p, &end);
    } else {
      v.kind = Setting::INT;
      v.i = strtoll(p, &end, 0);
      if (end && (*end == 'L' || *end == 'l')) end++;
      if (end && (*end == 'L' || *end == 'l')) end++;
    }
    if (!end || *end) return fail("bad number '" + tok + "'");
    return true;
  }
  bool group_body(Setting& g, char close) {
    for (;;) {
      skip();
      if (pos >= src.size()) {
        if (close == 0) return true;
        return fail("missing '}'");
      }
      if (src[pos] == close) {
        pos++;
        return true;
      }
      std::string n;
      if (!name(n)) return false;
      skip();
      if (pos < src.size() && (src[pos] == '=' || src[pos] == ':')) pos++;
      else return fail("expected '=' or ':' after " + n);
      auto child = std::make_unique<Setting>();
      if (!value(*child)) return false;
      skip();
      if (pos < src.size() && (src[pos] == ';' || src[pos] == ',')) pos++;
      g.children[n] = std::move(child);
    }
  }
};

const Setting* lookup(const Setting* g, const char* n) {
  if (!g || g->kind != Setting::GROUP) return nullptr;
  auto it = g->children.find(n);
  return it == g->children.end() ? nullptr : it->second.get();
}

}  // namespace

// readConfig::readConfig, src/independant/readConfig.cc:29-61
readConfig::readConfig(const char* configfile) {
  std::ifstream in(configfile ? configfile : "");
  if (!in) throw cdx::Error(CDX_ERR_CONFIG, std::string("cannot read config file ") + (configfile ? configfile : "(null)"));
  std::stringstream ss;
  ss << in.rdbuf();
  std::string text = ss.str();
  Setting root;
  Parser ps(text);
  if (!ps.group_body(root, 0)) throw cdx::Error(CDX_ERR_CONFIG, std::string(configfile) + ":" + ps.err);
  auto keep = [&](const std::string& s) {
    strings_.push_back(s);
    return strings_.back().c_str();
  };
  const Setting* images = lookup(&root, "InputImages");
  const Setting* def = lookup(images, "default");
  const Setting* pup = lookup(images, "pupil");
  auto str_into = [&](const Setting* g, const char* n, const char*& dst) {
    const Setting* s = lookup(g, n);
    if (s && s->kind == Setting::STRING) dst = keep(s->s);
  };
  str_into(def, "Intensity", common.Intensity);
  str_into(def, "Phase", common.Phase);
  str_into(def, "restart", common.restart);
  str_into(def, "Pattern", common.Pattern);
  str_into(pup, "Intensity", pupil.Intensity);
  str_into(pup, "Phase", pupil.Phase);
  str_into(pup, "restart", pupil.restart);
  str_into(pup, "Pattern", pupil.Pattern);
  const Setting* job = lookup(&root, "Job");
#define GETBOOL(x, y)                                  \
  if (const Setting* s = lookup(job, #x))             \
    if (s->kind == Setting::BOOL) x = s->b;
#define GETINT(x, y)                                   \
  if (const Setting* s = lookup(job, #x))             \
    if (s->kind == Setting::INT) x = (int)s->i;
#define GETREAL(x, y)                                  \
  if (const Setting* s = lookup(job, #x)) {           \
    if (s->kind == Setting::FLOAT) x = s->f;          \
    else if (s->kind == Setting::INT) x = (double)(int)s->i; \
  }
#define GETSTR(x, y)                                   \
  if (const Setting* s = lookup(job, #x))             \
    if (s->kind == Setting::STRING) x = keep(s->s);
  BOOLVAR(GETBOOL)
  INTVAR(GETINT)
  REALVAR(GETREAL)
  STRVAR(GETSTR)
#undef GETBOOL
#undef GETINT
#undef GETREAL
#undef GETSTR
}

void readConfig::print() {
  auto s = [](const char* p) { return p ? p : "(null)"; };
  printf("common Intensity=%s\ncommon Phase=%s\ncommon restart=%s\ncommon Pattern=%s\n", s(common.Intensity),
         s(common.Phase), s(common.restart), s(common.Pattern));
  printf("pupil Intensity=%s\npupil Phase=%s\npupil restart=%s\npupil Pattern=%s\n", s(pupil.Intensity),
         s(pupil.Phase), s(pupil.restart), s(pupil.Pattern));
#define PB(x, y) printf("bool: %s = %d  (default = %d)\n", #x, (int)x, (int)(y));
#define PI(x, y) printf("int: %s = %d  (default = %d)\n", #x, x, (int)(y));
#define PR(x, y) printf("float: %s = %g  (default = %g)\n", #x, x, (double)(y));
#define PS(x, y) printf("string: %s = %s  (default = %s)\n", #x, s(x), y);
  BOOLVAR(PB) INTVAR(PI) REALVAR(PR) STRVAR(PS)
#undef PB
#undef PI
#undef PR
#undef PS
}

// ---------------------------------------------------------------------------
// AlgoParser.  Parenthesised groups become sub-parsers (any nesting depth;
// the reference's search for the closing parenthesis is reliable for one
// level, where both agree).  next() is the reference's state machine,
// readConfig.cc:139-163.
// ---------------------------------------------------------------------------
namespace {
enum { kCnt = 10 };  // enum Algorithm::cnt, include/format.hpp:22
int token_id(const std::string& s) {
  static const std::pair<const char*, int> tab[] = {{"RAAR", CDX_RAAR}, {"HIO", CDX_HIO}, {"ER", CDX_ER},
                                                    {"POSER", CDX_POSER}, {"POSHIO", CDX_POSHIO},
                                                    {"FHIO", CDX_FHIO}, {"KKT", CDX_KKT},
                                                    {"shrinkWrap", CDX_SHRINKWRAP}, {"TV", CDX_TV},
                                                    {"XCORRELATION", CDX_XCORRELATION}};
  for (auto& e : tab)
    if (s == e.first) return e.second;
  return -1;
}
}  // namespace

#define SUBS (*(std::vector<AlgoParser*>*)subParsersp)
#define COUNT (*(std::vector<int>*)countp)
#define ALGOS (*(std::vector<int>*)algoListp)

AlgoParser::AlgoParser(const char* f) {
  nAlgo = kCnt;
  subParsersp = new std::vector<AlgoParser*>();
  countp = new std::vector<int>();
  algoListp = new std::vector<int>();
  std::string s;
  for (const char* p = f ? f : ""; *p; p++)
    if (*p != ' ') s += *p;
  size_t i = 0;
  auto bad = [&](const std::string& m) {
    throw cdx::Error(CDX_ERR_SCHEDULE, "algorithm \"" + std::string(f ? f : "") + "\": " + m);
  };
  while (i < s.size()) {
    size_t star = s.find('*', i);
    if (star == std::string::npos) bad("expected count*TOKEN");
    int num = atoi(s.substr(i, star - i).c_str());
    i = star + 1;
    int id;
    if (i < s.size() && s[i] == '(') {
      int depth = 0;
      size_t k = i;
      for (; k < s.size(); k++) {
        if (s[k] == '(') depth++;
        else if (s[k] == ')' && --depth == 0) break;
      }
      if (k >= s.size()) bad("unbalanced parenthesis");
      SUBS.push_back(new AlgoParser(s.substr(i + 1, k - i - 1).c_str()));
      id = nAlgo + (int)SUBS.size() - 1;
      i = k + 1;
      if (i < s.size() && s[i] != '+') bad("expected '+' after a group");
    } else {
      size_t k = s.find('+', i);
      if (k == std::string::npos) k = s.size();
      std::string tok = s.substr(i, k - i);
      id = token_id(tok);
      if (id < 0) bad("Algorithm " + tok + " not found");
      i = k;
    }
    COUNT.push_back(num);
    ALGOS.push_back(id);
    if (i < s.size() && s[i] == '+') i++;
  }
  if (ALGOS.empty()) bad("empty schedule");
  restart();
}

AlgoParser::~AlgoParser() {
  for (auto* p : SUBS) delete p;
  delete (std::vector<AlgoParser*>*)subParsersp;
  delete (std::vector<int>*)countp;
  delete (std::vector<int>*)algoListp;
}

void AlgoParser::restart() {
  currentAlgo = 0;
  currentCount = COUNT[0];
  for (auto* sub : SUBS) sub->restart();
}

int AlgoParser::next() {
  if (currentCount == 0) {
    if (currentAlgo == ALGOS.size() - 1) return -1;
    currentCount = COUNT[++currentAlgo];
  }
  if (ALGOS[currentAlgo] >= nAlgo) {
    AlgoParser* sub = SUBS[ALGOS[currentAlgo] - nAlgo];
    int r = sub->next();
    if (r == -1) {
      currentCount--;
      sub->restart();
      return next();
    }
    return r;
  }
  currentCount--;
  return ALGOS[currentAlgo];
}

extern "C" int cdx_expand_schedule(const char* algorithm, int* out, int cap) {
  try {
    AlgoParser p(algorithm);
    int n = 0, t;
    while ((t = p.next()) >= 0) {
      if (out && n < cap) out[n] = t;
      n++;
      if (n > 10000000) return -CDX_ERR_SCHEDULE;  // a zero count after the first element never ends
    }
    return n;
  } catch (const cdx::Error&) {
    return -CDX_ERR_SCHEDULE;
  }
}

// readConfig + print as JSON (host only): every field of the drop-in class.
extern "C" int cdx_read_config_json(const char* path, char* buf, int cap) {
  if (!buf || cap <= 0) {
    cdx::set_last_error("cdx_read_config_json: null buffer");
    return CDX_ERR_ARG;
  }
  try {
    readConfig c(path);
    std::string o = "{";
    auto esc = [](const char* p) {
      std::string r = "\"";
      for (; p && *p; p++) {
        if (*p == '"' || *p == '\\') r += '\\';
        r += *p;
      }
      return r + "\"";
    };
    auto str = [&](const char* p) { return p ? esc(p) : std::string("null"); };
    char num[64];
#define JB(x, y) o += "\"" #x "\": " + std::string(c.x ? "true" : "false") + ", ";
#define JI(x, y) o += "\"" #x "\": " + std::to_string(c.x) + ", ";
#define JR(x, y) snprintf(num, sizeof num, "%.17g", c.x); o += "\"" #x "\": " + std::string(num) + ", ";
#define JS(x, y) o += "\"" #x "\": " + str(c.x) + ", ";
    BOOLVAR(JB) INTVAR(JI) REALVAR(JR) STRVAR(JS)
#undef JB
#undef JI
#undef JR
#undef JS
    o += "\"common\": {\"Pattern\": " + str(c.common.Pattern) + ", \"Intensity\": " + str(c.common.Intensity) +
         ", \"Phase\": " + str(c.common.Phase) + ", \"restart\": " + str(c.common.restart) + "}, ";
    o += "\"pupil\": {\"Pattern\": " + str(c.pupil.Pattern) + ", \"Intensity\": " + str(c.pupil.Intensity) +
         ", \"Phase\": " + str(c.pupil.Phase) + ", \"restart\": " + str(c.pupil.restart) + "}}";
    if ((int)o.size() >= cap) {
      cdx::set_last_error("cdx_read_config_json: buffer too small");
      return CDX_ERR_ARG;
    }
    memcpy(buf, o.c_str(), o.size() + 1);
    return CDX_OK;
  } catch (const cdx::Error& e) {
    cdx::set_last_error(e.what());
    return e.code;
  }
}
