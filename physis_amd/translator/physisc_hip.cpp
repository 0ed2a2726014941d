// physisc-hip -- translates Physis DSL C programs to HIP for MI355X (gfx950).
//
// The HIP counterpart of the reference's `physisc --cuda` (translator/
// physisc.cc:449-690 driving cuda_translator.cc and cuda_runtime_builder.cc).
// ROSE is not available, so the front end is a restricted one for the DSL
// subset (docs/programming.md:139-194): the system C preprocessor expands the
// source, a tokenizer and brace matcher find the top-level functions, and
// the stencil kernels (the functions passed to PSStencilMap) are analysed
// token by token.
//
// Lowering (names follow the reference's generated code):
//   PSGrid<r>D<T>New(n..[, attr])   -> __PSGridNewHip(sizeof(T), r, attr, n..)
//                                      (reference_translator.cc:186-231)
//   PSStencilMap(k, dom, args..)     -> __PSStencilMap_k(dom, args..) returning
//                                      struct __PSStencil_k (reference_runtime_builder.cc:398-550)
//   PSStencilRun(maps..[, iter])     -> __PSStencilRun_<id>(iter, maps..), iter
//                                      defaulting to 1 (reference_translator.cc:1249-1283)
//   PSReduce(&v, op, g)              -> __PSReduceGridFloat/Double(&v, op, g)
//                                      (reference_translator.cc:1315-1351)
//   PSGridReduce(...)                -> the same as PSReduce(...) (an alias)
//   PSReduce(&v, op, k, dom, args..) -> __PSReduceMap_k(&v, op, dom, args..):
//                                      the reduction kernel k (returns float or
//                                      double) evaluated at every point of dom
//                                      and folded in FP64 (psk::RunReduceMap);
//                                      documented in the reference
//                                      (docs/programming.md, "Reduction") but
//                                      not implemented by its translators
//                                      (reference_translator.cc:1357-1360)
//   kernel k                         -> template <class SF> struct __PSPlan_k, a
//                                      Plan for the skeletons in include/physis/hip
//   PSGridGet(g, x+dx, y+dy, z+dz)   -> LDS/register tap for the staged grid,
//                                      else a load at the computed offset
//   PSGridGetPeriodic(...)           -> same with (i + n) % n wrapping
//   PSGridEmit(g, v)                 -> store at the current point
//   PSGridDim(g, d)                  -> g.dim[d]
// Stencil-range analysis (reference stencil_analysis.cc:248,
// runtime_builder.cc:78-118) picks the staged grid and its halo radius; a
// kernel that reads one float grid only at the 7-point star offsets and emits
// one value becomes a star plan (vectorised skeleton, and two maps of a
// ping-pong PSStencilRun fuse into one pass per two steps).
//
//   DeclareGrid<r>D(Name, struct T) -> the declaration plus T's layout
//                                      description and SoA plane helpers;
//                                      PSGridGet(g, ..).m[i], PSGridEmit(g, v)
//                                      and PSGridEmitUtype(g.m[i], v) address
//                                      one plane per scalar of T
//                                      (reference cuda_runtime_builder.cc:319-1102)
//   PSStencilMapRedBlack/Red/Black   -> the map run over the points of one or
//                                      both colours (x+y+z+c even)
//
// Not supported (rejected with a message): grids captured from globals,
// nested structs inside point types.
//
// Usage: physisc-hip [--config FILE] [-o OUT] [-I DIR]... [-D NAME[=V]]... INPUT.c
// Config keys (key = value lines): TRACE_KERNEL, KERNEL_ERROR_CHECK (bools),
// FUSE_STEPS (bool, default true).
#include <sys/stat.h>
#include <unistd.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void Die(const std::string &msg) {
  fprintf(stderr, "physisc-hip: error: %s\n", msg.c_str());
  exit(1);
}

// ---------------------------------------------------------------- tokens --
struct Tok {
  enum Kind { kIdent, kNum, kStr, kChar, kPunct } kind;
  std::string s;
};

bool IsIdentStart(char c) { return isalpha((unsigned char)c) || c == '_'; }
bool IsIdentChar(char c) { return isalnum((unsigned char)c) || c == '_'; }

std::vector<Tok> Tokenize(const std::string &src) {
  static const char *kPuncts[] = {"<<=", ">>=", "...", "->", "++", "--", "<<", ">>", "<=", ">=", "==", "!=",
                                  "&&",  "||",  "+=",  "-=", "*=", "/=", "%=", "&=", "|=", "^=", "##", "::"};
  std::vector<Tok> out;
  size_t i = 0, n = src.size();
  while (i < n) {
    char c = src[i];
    if (isspace((unsigned char)c)) {
      ++i;
      continue;
    }
    if (c == '/' && i + 1 < n && src[i + 1] == '/') {
      while (i < n && src[i] != '\n') ++i;
      continue;
    }
    if (c == '/' && i + 1 < n && src[i + 1] == '*') {
      size_t e = src.find("*/", i + 2);
      i = e == std::string::npos ? n : e + 2;
      continue;
    }
    if (c == '#') {  // leftover directive (e.g. #pragma): keep the line verbatim
      size_t e = src.find('\n', i);
      if (e == std::string::npos) e = n;
      out.push_back({Tok::kPunct, src.substr(i, e - i) + "\n"});
      i = e;
      continue;
    }
    if (IsIdentStart(c)) {
      size_t j = i;
      while (j < n && IsIdentChar(src[j])) ++j;
      out.push_back({Tok::kIdent, src.substr(i, j - i)});
      i = j;
      continue;
    }
    if (isdigit((unsigned char)c) || (c == '.' && i + 1 < n && isdigit((unsigned char)src[i + 1]))) {
      size_t j = i;
      while (j < n && (IsIdentChar(src[j]) || src[j] == '.' ||
                       ((src[j] == '+' || src[j] == '-') && (src[j - 1] == 'e' || src[j - 1] == 'E') &&
                        !(src[i] == '0' && i + 1 < n && (src[i + 1] == 'x' || src[i + 1] == 'X')))))
        ++j;
      out.push_back({Tok::kNum, src.substr(i, j - i)});
      i = j;
      continue;
    }
    if (c == '"' || c == '\'') {
      size_t j = i + 1;
      while (j < n && src[j] != c) j += (src[j] == '\\') ? 2 : 1;
      out.push_back({c == '"' ? Tok::kStr : Tok::kChar, src.substr(i, j + 1 - i)});
      i = j + 1;
      continue;
    }
    bool matched = false;
    for (const char *p : kPuncts) {
      size_t l = strlen(p);
      if (src.compare(i, l, p) == 0) {
        out.push_back({Tok::kPunct, p});
        i += l;
        matched = true;
        break;
      }
    }
    if (!matched) {
      out.push_back({Tok::kPunct, std::string(1, c)});
      ++i;
    }
  }
  return out;
}

size_t MatchClose(const std::vector<Tok> &t, size_t open) {
  const std::string o = t[open].s;
  const std::string c = o == "(" ? ")" : o == "[" ? "]" : "}";
  int depth = 0;
  for (size_t i = open; i < t.size(); ++i) {
    if (t[i].kind != Tok::kPunct) continue;
    if (t[i].s == o) ++depth;
    else if (t[i].s == c && --depth == 0) return i;
  }
  Die("unbalanced '" + o + "'");
}

// Splits (b, e) at top-level commas into [begin, end) ranges.
std::vector<std::pair<size_t, size_t>> SplitArgs(const std::vector<Tok> &t, size_t b, size_t e) {
  std::vector<std::pair<size_t, size_t>> out;
  if (b >= e) return out;
  int depth = 0;
  size_t start = b;
  for (size_t i = b; i < e; ++i) {
    if (t[i].kind != Tok::kPunct) continue;
    const std::string &s = t[i].s;
    if (s == "(" || s == "[" || s == "{") ++depth;
    else if (s == ")" || s == "]" || s == "}") --depth;
    else if (s == "," && depth == 0) {
      out.push_back({start, i});
      start = i + 1;
    }
  }
  out.push_back({start, e});
  return out;
}

std::string Text(const std::vector<Tok> &t, size_t b, size_t e) {
  std::string out;
  for (size_t i = b; i < e; ++i) {
    if (!out.empty()) {
      char p = out.back();
      bool need = (IsIdentChar(p) || p == '.') && (IsIdentChar(t[i].s[0]) || t[i].s[0] == '.');
      if (need || (t[i].kind != Tok::kPunct) || (t[i].s != ")" && t[i].s != "," && t[i].s != ";" &&
                                                 t[i].s != "]" && p != '(' && p != '['))
        out += ' ';
    }
    out += t[i].s;
    if (t[i].s.size() > 1 && t[i].s[0] == '#') out += "";
  }
  return out;
}

// ------------------------------------------------------------ grid types --
struct GridType {
  int rank = 0;
  std::string elem;   // float double int long, or the C type of a point struct
  std::string utype;  // name of a user-defined point type ("" for primitives)
};

// A user-defined point type: DeclareGrid<r>D(Name, struct T) with T a struct
// of scalars and arrays of scalars (reference docs/programming.md:118-137).
struct UMember {
  std::string name;
  std::string ctype;      // scalar type text
  std::vector<int> dims;  // array extents (empty for a scalar member)
  int count = 1;          // product of dims
};

struct UType {
  std::string name;   // Name of PSGrid<r>D<Name>
  std::string ctype;  // "struct T" or a typedef name
  int rank = 0;
  std::vector<UMember> members;
  const UMember *Find(const std::string &m) const {
    for (const UMember &u : members)
      if (u.name == m) return &u;
    return nullptr;
  }
};

// Grid type name (PSGrid3DPoint) -> point type, filled before any parsing
// that needs grid types.
std::map<std::string, UType> g_utypes;

bool ParseGridType(const std::string &id, GridType *gt) {
  // PSGrid<r>D<Float|Double|Int|Long>, or a declared PSGrid<r>D<Name>
  if (id.size() < 9 || id.compare(0, 6, "PSGrid") != 0 || !isdigit((unsigned char)id[6]) || id[7] != 'D')
    return false;
  std::string rest = id.substr(8);
  static const std::map<std::string, std::string> kElem = {
      {"Float", "float"}, {"Double", "double"}, {"Int", "int"}, {"Long", "long"}};
  auto it = kElem.find(rest);
  if (it == kElem.end()) {
    auto u = g_utypes.find(id);
    if (u == g_utypes.end()) return false;
    gt->rank = u->second.rank;
    gt->elem = u->second.ctype;
    gt->utype = u->second.name;
    return true;
  }
  gt->rank = id[6] - '0';
  gt->elem = it->second;
  gt->utype.clear();
  return gt->rank >= 1 && gt->rank <= 3;
}

// Index of the '{' that matches the '}' at `close`.
size_t MatchOpen(const std::vector<Tok> &t, size_t close) {
  int depth = 0;
  for (size_t i = close + 1; i-- > 0;) {
    if (t[i].kind != Tok::kPunct) continue;
    if (t[i].s == "}") ++depth;
    else if (t[i].s == "{" && --depth == 0) return i;
  }
  Die("unbalanced '}'");
}

// Members of the struct body between braces lb and rb: scalar members and
// arrays of scalars with constant extents (macros are already expanded).
std::vector<UMember> ParseStructBody(const std::vector<Tok> &t, size_t lb, size_t rb, const std::string &what) {
  static const std::set<std::string> kTypeWords = {"float", "double", "int",      "long",  "short",
                                                   "char",  "unsigned", "signed", "const"};
  std::vector<UMember> out;
  size_t i = lb + 1;
  while (i < rb) {
    size_t semi = i;
    while (semi < rb && t[semi].s != ";") {
      if (t[semi].s == "{") Die(what + ": nested aggregates are not supported as point members");
      ++semi;
    }
    if (semi == rb) Die(what + ": expected ';' after a member");
    std::string ctype;
    size_t q = i;
    while (q < semi && t[q].kind == Tok::kIdent && kTypeWords.count(t[q].s)) {
      if (t[q].s != "const") ctype += (ctype.empty() ? "" : " ") + t[q].s;
      ++q;
    }
    if (ctype.empty()) Die(what + ": point members must have scalar C types");
    for (auto &d : SplitArgs(t, q, semi)) {
      size_t a = d.first;
      if (a >= d.second || t[a].kind != Tok::kIdent) Die(what + ": unsupported member declarator");
      UMember m;
      m.name = t[a].s;
      m.ctype = ctype;
      size_t b = a + 1;
      while (b < d.second && t[b].s == "[") {
        size_t c = MatchClose(t, b);
        char *end = nullptr;
        long v = c == b + 2 && t[b + 1].kind == Tok::kNum ? strtol(t[b + 1].s.c_str(), &end, 0) : 0;
        if (v <= 0 || !end || *end != '\0') Die(what + ": array extents of members must be integer constants");
        m.dims.push_back((int)v);
        m.count *= (int)v;
        b = c + 1;
      }
      if (b != d.second) Die(what + ": unsupported member declarator");
      out.push_back(m);
    }
    i = semi + 1;
  }
  if (out.empty()) Die(what + ": point type without members");
  return out;
}

// Finds every DeclareGrid<r>D(Name, type) and the definition of its struct.
void CollectUtypes(const std::vector<Tok> &t) {
  for (size_t i = 0; i + 1 < t.size(); ++i) {
    const std::string &id = t[i].s;
    if (t[i].kind != Tok::kIdent || t[i + 1].s != "(" || id.size() != 13 || id.compare(0, 11, "DeclareGrid") != 0 ||
        id[12] != 'D' || id[11] < '1' || id[11] > '3')
      continue;
    size_t close = MatchClose(t, i + 1);
    auto args = SplitArgs(t, i + 2, close);
    if (args.size() != 2 || args[0].second != args[0].first + 1) Die(id + ": expected (Name, type)");
    UType u;
    u.name = t[args[0].first].s;
    u.rank = id[11] - '0';
    u.ctype = Text(t, args[1].first, args[1].second);
    static const std::set<std::string> kPrimitive = {"float", "double", "int", "long"};
    if (kPrimitive.count(u.ctype)) continue;  // the built-in grid types
    size_t lb = 0, rb = 0;
    const size_t tb = args[1].first, te = args[1].second;
    if (te - tb == 2 && t[tb].s == "struct") {
      const std::string &tag = t[tb + 1].s;
      for (size_t j = 0; j + 2 < i; ++j)
        if (t[j].s == "struct" && t[j + 1].s == tag && t[j + 2].s == "{") {
          lb = j + 2;
          rb = MatchClose(t, lb);
          break;
        }
    } else if (te - tb == 1) {
      const std::string &name = t[tb].s;  // typedef struct [tag] { ... } name;
      for (size_t j = 1; j + 1 < i; ++j)
        if (t[j].s == name && t[j - 1].s == "}" && t[j + 1].s == ";") {
          rb = j - 1;
          lb = MatchOpen(t, rb);
          break;
        }
    }
    if (!rb) Die(id + "(" + u.name + ", " + u.ctype + "): the point struct must be defined before it");
    u.members = ParseStructBody(t, lb, rb, "point type " + u.ctype);
    g_utypes["PSGrid" + std::string(1, id[11]) + "D" + u.name] = u;
  }
}

bool ParseGridNew(const std::string &id, GridType *gt) {
  if (id.size() < 4 || id.compare(id.size() - 3, 3, "New") != 0) return false;
  return ParseGridType(id.substr(0, id.size() - 3), gt);
}

std::string StorageOf(const GridType &g) {
  return g.elem == "float" ? "SF" : "psk::Native<" + g.elem + ">";
}

// Linear index of a member element: ((s0) * d1 + (s1)) * d2 + (s2) ...
std::string MemberIndex(const UMember &m, const std::vector<std::string> &subs) {
  std::string lin = "(" + subs[0] + ")";
  for (size_t d = 1; d < subs.size(); ++d)
    lin = "(" + lin + " * " + std::to_string(m.dims[d]) + " + (" + subs[d] + "))";
  return lin;
}

// Support code for one point type, emitted right after its DeclareGrid:
// the layout description, the devNew/devFree/devCopyin/devCopyout functions
// __PSGridNew & co. take, the kernel-side reference (one pointer per
// member's first plane) and whole-point load/store helpers.
std::string EmitUtypeSupport(const UType &u) {
  std::ostringstream o;
  const std::string &N = u.name, &T = u.ctype;
  const size_t nm = u.members.size();
  o << "\n// Point type " << T << " (grid type PSGrid" << u.rank << "D" << N << "): " << nm
    << " members, one SoA plane per scalar.\n";
  o << "static const int __PSUOff_" << N << "[] = {";
  for (size_t m = 0; m < nm; ++m) o << (m ? ", " : "") << "(int)offsetof(" << T << ", " << u.members[m].name << ")";
  o << "};\nstatic const int __PSUSize_" << N << "[] = {";
  for (size_t m = 0; m < nm; ++m) o << (m ? ", " : "") << "(int)sizeof(" << u.members[m].ctype << ")";
  o << "};\nstatic const int __PSUCount_" << N << "[] = {";
  for (size_t m = 0; m < nm; ++m) o << (m ? ", " : "") << u.members[m].count;
  o << "};\n";
  o << "static const PSHipUtype __PSUtype_" << N << " = {(int)sizeof(" << T << "), " << nm << ", __PSUOff_" << N
    << ", __PSUSize_" << N << ", __PSUCount_" << N << "};\n";
  o << "static void *__PSGridDevNew_" << N << "(int nd, PSVectorInt dim) { return __PSHipUtypeDevNew(&__PSUtype_"
    << N << ", nd, dim); }\n";
  o << "static void __PSGridDevFree_" << N << "(void *d) { __PSHipUtypeDevFree(d); }\n";
  o << "static void __PSGridDevCopyin_" << N << "(void *d, const void *src) { __PSHipUtypeCopyin(d, src); }\n";
  o << "static void __PSGridDevCopyout_" << N << "(void *d, void *dst) { __PSHipUtypeCopyout(d, dst); }\n";
  o << "static __PSGrid *__PSGridNewUtype_" << N << "(int d0, int d1, int d2) {\n"
    << "  PSVectorInt d = {d0, d1, d2};\n  (void)__PSGridDevFree_" << N << ";\n  (void)__PSGridDevCopyin_" << N
    << ";\n  (void)__PSGridDevCopyout_" << N << ";\n"
    << "  return __PSGridNew((int)sizeof(" << T << "), " << u.rank << ", d, __PSGridDevNew_" << N << ");\n}\n";
  o << "struct __PSURef_" << N << " {\n  int dim[3];\n  int64_t ps;  // elements between planes\n";
  for (const UMember &m : u.members) o << "  " << m.ctype << " *" << m.name << ";\n";
  o << "};\n";
  o << "static inline struct __PSURef_" << N << " __PSUMakeRef_" << N << "(const __PSGrid *g) {\n"
    << "  struct __PSURef_" << N << " r;\n  const __PSGridUtype_dev *d = (const __PSGridUtype_dev *)g->dev;\n"
    << "  for (int i = 0; i < 3; ++i) r.dim[i] = d->dim[i];\n  r.ps = d->plane_stride;\n";
  for (size_t m = 0; m < nm; ++m)
    o << "  r." << u.members[m].name << " = (" << u.members[m].ctype << " *)__PSHipUtypeMember(g, " << m << ");\n";
  o << "  return r;\n}\n";
  o << "__device__ __forceinline__ " << T << " __PSUGet_" << N << "(const struct __PSURef_" << N
    << " &g, int64_t off) {\n  " << T << " v;\n";
  for (const UMember &m : u.members) {
    if (m.dims.empty()) o << "  v." << m.name << " = g." << m.name << "[off];\n";
    else
      o << "  for (int c = 0; c < " << m.count << "; ++c) ((" << m.ctype << " *)v." << m.name << ")[c] = g."
        << m.name << "[c * g.ps + off];\n";
  }
  o << "  return v;\n}\n";
  o << "__device__ __forceinline__ void __PSUPut_" << N << "(const struct __PSURef_" << N << " &g, int64_t off, const "
    << T << " &v) {\n";
  for (const UMember &m : u.members) {
    if (m.dims.empty()) o << "  g." << m.name << "[off] = v." << m.name << ";\n";
    else
      o << "  for (int c = 0; c < " << m.count << "; ++c) g." << m.name << "[c * g.ps + off] = ((const " << m.ctype
        << " *)v." << m.name << ")[c];\n";
  }
  o << "}\n";
  return o.str();
}

// ------------------------------------------------------------- functions --
struct Param {
  std::string type;  // text of the type tokens
  std::string name;
  bool is_grid = false;
  GridType gt;
};

struct Func {
  std::string name;
  size_t begin;   // first token of the definition (return type)
  size_t lparen, rparen, lbrace, rbrace;
  std::vector<Param> params;
};

std::vector<Func> FindFunctions(const std::vector<Tok> &t) {
  std::vector<Func> out;
  size_t stmt_begin = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    const Tok &k = t[i];
    if (k.kind == Tok::kPunct && (k.s == ";" || (k.s.size() > 1 && k.s[0] == '#'))) {
      stmt_begin = i + 1;
      continue;
    }
    if (k.kind == Tok::kPunct && k.s == "{" && i >= 2 && t[i - 1].kind == Tok::kStr && t[i - 2].s == "extern") {
      stmt_begin = i + 1;  // extern "C" { ... }: its definitions are top level
      continue;
    }
    if (k.kind == Tok::kPunct && k.s == "}") {  // end of an extern "C" block
      stmt_begin = i + 1;
      continue;
    }
    if (k.kind == Tok::kPunct && k.s == "{") {
      // A top-level brace: function body if preceded by ") " (with optional
      // attributes), else a struct/union/enum/initializer body.
      size_t close = MatchClose(t, i);
      size_t j = i;
      while (j > stmt_begin && t[j - 1].kind == Tok::kIdent &&
             (t[j - 1].s == "const" || t[j - 1].s == "__attribute__"))
        --j;
      if (j > stmt_begin && t[j - 1].s == ")") {
        // find the matching '(' of the parameter list
        int depth = 0;
        size_t rp = j - 1, lp = rp;
        for (size_t q = rp + 1; q-- > stmt_begin;) {
          if (t[q].s == ")") ++depth;
          else if (t[q].s == "(" && --depth == 0) {
            lp = q;
            break;
          }
        }
        if (lp > stmt_begin && t[lp - 1].kind == Tok::kIdent) {
          Func f;
          f.name = t[lp - 1].s;
          f.begin = stmt_begin;
          f.lparen = lp;
          f.rparen = rp;
          f.lbrace = i;
          f.rbrace = close;
          for (auto &a : SplitArgs(t, lp + 1, rp)) {
            if (a.first >= a.second) continue;
            if (a.second - a.first == 1 && t[a.first].s == "void") continue;
            Param p;
            size_t name_at = a.second - 1;
            while (name_at > a.first && t[name_at].kind != Tok::kIdent) --name_at;  // skip [] etc.
            p.name = t[name_at].s;
            p.type = Text(t, a.first, name_at);
            for (size_t q = a.first; q < name_at; ++q)
              if (ParseGridType(t[q].s, &p.gt)) p.is_grid = true;
            f.params.push_back(p);
          }
          out.push_back(f);
        }
      }
      i = close;
      // A struct definition continues until ';'; a function ends here.
      if (!out.empty() && out.back().rbrace == close) stmt_begin = close + 1;
      continue;
    }
  }
  return out;
}

// ----------------------------------------------------------------- config --
struct Config {
  bool trace = false;
  bool error_check = false;
  bool fuse = true;
};

Config LoadConfig(const std::string &path) {
  Config c;
  if (path.empty()) return c;
  std::ifstream in(path);
  if (!in) Die("cannot read config " + path);
  std::string line;
  while (std::getline(in, line)) {
    size_t h = line.find_first_of("#-");
    if (h != std::string::npos && (line[h] == '#' || (h + 1 < line.size() && line[h + 1] == '-')))
      line = line.substr(0, h);
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      size_t b = s.find_first_not_of(" \t\r"), e = s.find_last_not_of(" \t\r,;");
      return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
    };
    std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
    bool on = v == "true" || v == "1" || v == "TRUE" || v == "True";
    if (k == "TRACE_KERNEL") c.trace = on;
    else if (k == "KERNEL_ERROR_CHECK" || k == "CUDA_KERNEL_ERROR_CHECK") c.error_check = on;
    else if (k == "FUSE_STEPS") c.fuse = on;
  }
  return c;
}

// ----------------------------------------------------------------- kernels --
struct Get {
  std::string grid;
  bool periodic = false;
  bool relative = false;
  int off[3] = {0, 0, 0};
};

struct Kernel {
  const Func *f = nullptr;
  int rank = 0;
  std::vector<std::string> idx;           // index parameter names
  std::vector<Param> grids, scalars;       // in parameter order
  std::vector<Param> all;                  // all non-index params in order
  std::map<std::string, GridType> gtype;
  std::vector<Get> gets;
  std::set<std::string> emitted;
  int emit_count = 0;
  std::string staged;   // grid staged in LDS ("" = none)
  int rx = 0, ry = 0, rz = 0;
  bool wrap = false;
  bool star = false;    // vectorised star plan (compute() available)
  std::string out_grid; // the single emitted grid when star
  // Two ping-pong maps of a star kernel may fuse into one pass only when the
  // body never reads the grid it emits: in a fused pass the second map's
  // reads of its output grid (the first map's input) would see neither the
  // value the separate steps see nor a finished buffer.
  bool fusable = false;
  // Point form (psk::PointTB, skeleton_point_tb.h): a staged float grid read
  // within radius 1 on every axis (diagonals included), the other grids read
  // at the point itself only (point_grids, in parameter order), one float
  // grid emitted as the last statement, the index names used only inside
  // grid reads.  Two maps of such a kernel over the interior of the staged
  // grid run two sweeps per HBM pass.
  bool point = false;
  std::vector<std::string> point_grids;
  // Locals assigned once from PSGridDim(staged grid, d): name -> d.
  std::map<std::string, int> dimvar;
  int clamp_ifs = 0;    // clamped-face reads of the staged grid (ClampIf)
};

// The grid whose extents index every rank-3 grid of a 3-D kernel launched
// with uniform extents (psk::Get3/Put3 with __kU): the staged grid, else the
// first rank-3 grid of primitive type; "" when the kernel has none (or is
// not 3-D), and then its accesses keep per-grid offsets.
// A parameter type without a leading "const" (a struct member assigned
// after construction).
std::string StripConst(std::string t) {
  while (t.compare(0, 6, "const ") == 0) t = t.substr(6);
  return t;
}

std::string UniformRef(const Kernel &k) {
  if (k.rank != 3) return "";
  if (!k.staged.empty()) {
    const GridType &sg = k.gtype.at(k.staged);
    if (sg.rank == 3 && sg.utype.empty()) return k.staged;
  }
  for (const Param &p : k.all)
    if (p.is_grid && p.gt.rank == 3 && p.gt.utype.empty()) return p.name;
  return "";
}

// Parses "x", "x + 3", "x - 1", "2 + x", "(x-1)" against index name `name`.
bool RelativeIndex(const std::vector<Tok> &t, size_t b, size_t e, const std::string &name, int *off) {
  while (e - b >= 2 && t[b].s == "(" && MatchClose(t, b) == e - 1) {
    ++b;
    --e;
  }
  if (e - b == 1 && t[b].kind == Tok::kIdent && t[b].s == name) {
    *off = 0;
    return true;
  }
  auto num = [&](size_t i, int *v) {
    if (t[i].kind != Tok::kNum) return false;
    char *end = nullptr;
    long x = strtol(t[i].s.c_str(), &end, 0);
    if (*end != '\0') return false;
    *v = (int)x;
    return true;
  };
  int v;
  if (e - b == 3 && t[b].kind == Tok::kIdent && t[b].s == name && (t[b + 1].s == "+" || t[b + 1].s == "-") &&
      num(b + 2, &v)) {
    *off = t[b + 1].s == "+" ? v : -v;
    return true;
  }
  if (e - b == 3 && num(b, &v) && t[b + 1].s == "+" && t[b + 2].kind == Tok::kIdent && t[b + 2].s == name) {
    *off = v;
    return true;
  }
  return false;
}

class Translator {
 public:
  Translator(std::vector<Tok> toks, Config cfg, std::string input)
      : t_(std::move(toks)), cfg_(cfg), input_(std::move(input)) {}

  std::string Run(const std::vector<std::string> &includes);

 private:
  void CollectGridVars();
  void AnalyzeKernel(Kernel &k);
  std::string RewriteBody(const Kernel &k, size_t b, size_t e, bool compute_form, bool exact = false);
  // Whether the body's one PSGridEmit sits at its top level, followed only
  // by an optional "return;", with no return before it.
  bool EmitIsTail(const Func &f) const;
  // Set while RewriteBody renders a point-form update(): staged-grid reads
  // become __PSat(dx, dy, dz) and point-grid reads __PSpt(n).
  bool point_mode_ = false;
  int ClampIf(const Kernel &k, size_t i, size_t e) const;
  std::string EmitPlan(const Kernel &k);
  std::string EmitMapAndLaunch(const Kernel &k);
  // Locals bound from a plan's Args at the top of a body.
  std::string Prologue(const Kernel &k) const;
  // A reduction kernel (the third argument of a PSReduce with a domain):
  // its value plan and the host function the PSReduce call becomes.
  std::string EmitReduceKernel(const Kernel &k);
  std::string RewriteHost(size_t b, size_t e);
  // modes[i]: 0 plain map, 1 red-black (red then black pass), 2 red only, 3 black only.
  std::string EmitRun(const std::vector<std::pair<std::string, std::string>> &maps, const std::vector<int> &modes,
                      int id);

  std::vector<Tok> t_;
  Config cfg_;
  std::string input_;
  std::vector<Func> funcs_;
  std::map<std::string, Kernel> kernels_;
  std::map<std::string, Kernel> rkernels_;  // reduction kernels
  std::map<std::string, GridType> gridvars_;
  std::vector<std::string> runs_;  // emitted run functions
  int run_id_ = 0;
};

void Translator::CollectGridVars() {
  for (size_t i = 0; i + 1 < t_.size(); ++i) {
    GridType gt;
    if (t_[i].kind == Tok::kIdent && ParseGridType(t_[i].s, &gt)) {
      size_t j = i + 1;
      while (j < t_.size() && (t_[j].s == "*" || t_[j].s == "const")) ++j;
      while (j < t_.size() && t_[j].kind == Tok::kIdent) {
        gridvars_[t_[j].s] = gt;
        // more declarators: ", name"
        size_t q = j + 1;
        if (q < t_.size() && t_[q].s == "=") {
          int depth = 0;
          while (q < t_.size()) {
            const std::string &s = t_[q].s;
            if (s == "(" || s == "[" || s == "{") ++depth;
            else if (s == ")" || s == "]" || s == "}") { if (--depth < 0) break; }
            else if ((s == "," || s == ";") && depth == 0) break;
            ++q;
          }
        }
        if (q < t_.size() && t_[q].s == "," && q + 1 < t_.size() && t_[q + 1].kind == Tok::kIdent) {
          j = q + 1;
          continue;
        }
        break;
      }
    }
  }
}

void Translator::AnalyzeKernel(Kernel &k) {
  const Func &f = *k.f;
  size_t p = 0;
  for (; p < f.params.size(); ++p) {
    const Param &pr = f.params[p];
    if (pr.is_grid) break;
    if (pr.type.find("int") == std::string::npos) break;
    k.idx.push_back(pr.name);
  }
  k.rank = (int)k.idx.size();
  if (k.rank < 1 || k.rank > 3) Die("kernel " + f.name + ": expected 1-3 leading int index parameters");
  for (; p < f.params.size(); ++p) {
    const Param &pr = f.params[p];
    k.all.push_back(pr);
    if (pr.is_grid) {
      k.grids.push_back(pr);
      k.gtype[pr.name] = pr.gt;
    } else {
      if (pr.type.find("PSGrid") != std::string::npos)
        Die("kernel " + f.name + ": user-defined point types are not supported by this target");
      k.scalars.push_back(pr);
    }
  }
  // Scan accesses.
  std::map<std::string, int> nonzero;  // gets at non-zero offsets per grid
  std::map<std::string, bool> any_abs;
  for (size_t i = f.lbrace; i < f.rbrace; ++i) {
    if (t_[i].kind != Tok::kIdent || i + 1 >= f.rbrace || t_[i + 1].s != "(") continue;
    const std::string &name = t_[i].s;
    if (name == "PSGridGet" || name == "PSGridGetPeriodic") {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      if (args.empty()) Die("kernel " + f.name + ": empty PSGridGet");
      std::string g = Text(t_, args[0].first, args[0].second);
      auto it = k.gtype.find(g);
      if (it == k.gtype.end()) Die("kernel " + f.name + ": PSGridGet on unknown grid '" + g + "'");
      if ((int)args.size() != it->second.rank + 1)
        Die("kernel " + f.name + ": PSGridGet on '" + g + "' has the wrong number of indices");
      Get acc;
      acc.grid = g;
      acc.periodic = name == "PSGridGetPeriodic";
      acc.relative = it->second.rank <= k.rank;
      for (int d = 0; d < it->second.rank && acc.relative; ++d)
        acc.relative = RelativeIndex(t_, args[d + 1].first, args[d + 1].second, k.idx[d], &acc.off[d]);
      k.gets.push_back(acc);
      if (!acc.relative) any_abs[g] = true;
    } else if (name == "PSGridEmit") {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      std::string g = Text(t_, args[0].first, args[0].second);
      if (!k.gtype.count(g)) Die("kernel " + f.name + ": PSGridEmit on unknown grid '" + g + "'");
      k.emitted.insert(g);
      ++k.emit_count;
    } else if (name == "PSGridEmitUtype") {
      // PSGridEmitUtype(g.member[i].., v): one member of a point-type grid.
      std::string g = t_[i + 2].s;
      auto it = k.gtype.find(g);
      if (it == k.gtype.end() || it->second.utype.empty())
        Die("kernel " + f.name + ": PSGridEmitUtype on '" + g + "', which is not a grid of a point type");
      k.emitted.insert(g);
      ++k.emit_count;
    } else if (name == "PSGridSet") {
      Die("kernel " + f.name + ": " + name + " is not supported by this target");
    }
  }
  // Staging: the grid of the kernel's rank and element type read at the most
  // non-zero relative offsets (radius <= 2, not emitted).
  std::map<std::string, int> count;
  for (const Get &g : k.gets) {
    if (!g.relative) continue;
    const GridType &gt = k.gtype[g.grid];
    if (gt.rank != k.rank || k.emitted.count(g.grid) || !gt.utype.empty()) continue;
    if (g.off[0] || g.off[1] || g.off[2]) count[g.grid]++;
  }
  int best = 0;
  for (auto &c : count) {
    int r = 0;
    for (const Get &g : k.gets)
      if (g.grid == c.first && g.relative)
        for (int d = 0; d < 3; ++d) r = std::max(r, std::abs(g.off[d]));
    if (r <= 2 && c.second > best) {
      best = c.second;
      k.staged = c.first;
    }
  }
  if (!k.staged.empty()) {
    for (const Get &g : k.gets) {
      if (g.grid != k.staged || !g.relative) continue;
      k.rx = std::max(k.rx, std::abs(g.off[0]));
      k.ry = std::max(k.ry, std::abs(g.off[1]));
      k.rz = std::max(k.rz, std::abs(g.off[2]));
      if (g.periodic) k.wrap = true;
    }
    if (k.rank < 3) k.rz = 0;
    if (k.rank < 2) k.ry = 0;
  }
  // Star form: rank 3, staged float grid with offsets in the 7-point star,
  // no absolute reads of it, one unconditional emit as the last statement.
  if (k.rank == 3 && !k.staged.empty() && k.gtype[k.staged].elem == "float" && !any_abs[k.staged] &&
      k.emitted.size() == 1 && k.emit_count == 1) {
    bool in_star = true;
    for (const Get &g : k.gets) {
      if (g.grid != k.staged) continue;
      int nz = (g.off[0] != 0) + (g.off[1] != 0) + (g.off[2] != 0);
      if (nz > 1 || std::abs(g.off[0]) > 1 || std::abs(g.off[1]) > 1 || std::abs(g.off[2]) > 1) in_star = false;
    }
    const bool tail_ok = EmitIsTail(f);
    const GridType &og = k.gtype[*k.emitted.begin()];
    if (in_star && tail_ok && (k.rx | k.ry | k.rz) <= 1 && og.elem == "float" && og.rank == 3) {
      k.star = true;
      k.out_grid = *k.emitted.begin();
      k.fusable = true;
      for (const Get &g : k.gets)
        if (g.grid == k.out_grid) k.fusable = false;
    }
  }
  if (!k.star && k.rank == 3 && !k.staged.empty() && k.gtype[k.staged].elem == "float" &&
      k.gtype[k.staged].utype.empty() && k.gtype[k.staged].rank == 3 && !any_abs[k.staged] && !k.wrap &&
      (k.rx | k.ry | k.rz) <= 1 && k.emitted.size() == 1 && k.emit_count == 1 && EmitIsTail(f)) {
    const std::string out = *k.emitted.begin();
    const GridType &og = k.gtype[out];
    bool ok = out != k.staged && og.elem == "float" && og.rank == 3 && og.utype.empty();
    std::set<std::string> point;
    for (const Get &g : k.gets) {
      if (g.grid == out || g.periodic) ok = false;
      if (g.grid == k.staged) continue;
      const GridType &gt = k.gtype[g.grid];
      if (!g.relative || g.off[0] || g.off[1] || g.off[2] || gt.rank != 3 || gt.elem != "float" ||
          !gt.utype.empty())
        ok = false;
      point.insert(g.grid);
    }
    // The index names may appear only inside grid reads (update() has no
    // point coordinates), and PSGridDim not at all.
    for (size_t i = f.lbrace + 1; i < f.rbrace && ok; ++i) {
      const std::string &s = t_[i].s;
      if ((s == "PSGridGet" || s == "PSGridGetPeriodic") && i + 1 < f.rbrace && t_[i + 1].s == "(") {
        i = MatchClose(t_, i + 1);
        continue;
      }
      if (s == "PSGridDim") ok = false;
      for (const std::string &n : k.idx)
        if (t_[i].kind == Tok::kIdent && s == n) ok = false;
    }
    for (const Param &p : k.all)
      if (p.is_grid && point.count(p.name)) k.point_grids.push_back(p.name);
    if (ok && !k.point_grids.empty() && (int)k.point_grids.size() <= 16) {
      k.point = true;
      k.out_grid = out;
    } else {
      k.point_grids.clear();
    }
  }
  if (k.star && !k.wrap) {
    // Locals that hold a dimension of the staged grid: "n = PSGridDim(g, d);"
    // as the only assignment to n.
    // Writes to a local: any assignment operator or ++/-- next to it, or its
    // address taken.
    auto is_assign = [](const std::string &op) {
      return op == "++" || op == "--" ||
             (!op.empty() && op.back() == '=' && op != "==" && op != "<=" && op != ">=" && op != "!=");
    };
    std::map<std::string, int> assigned;
    for (size_t i = f.lbrace + 1; i < f.rbrace; ++i) {
      if (t_[i].kind != Tok::kIdent) continue;
      const bool after = i + 1 < f.rbrace && is_assign(t_[i + 1].s);
      const bool before = i > f.lbrace + 1 && (t_[i - 1].s == "++" || t_[i - 1].s == "--" || t_[i - 1].s == "&");
      if (after || before) ++assigned[t_[i].s];
    }
    for (size_t i = f.lbrace + 1; i + 9 < f.rbrace; ++i) {
      if (t_[i].kind == Tok::kIdent && t_[i + 1].s == "=" && t_[i + 2].s == "PSGridDim" && t_[i + 3].s == "(" &&
          t_[i + 4].s == k.staged && t_[i + 5].s == "," && t_[i + 6].kind == Tok::kNum && t_[i + 7].s == ")" &&
          t_[i + 8].s == ";" && assigned[t_[i].s] == 1) {
        const int d = atoi(t_[i + 6].s.c_str());
        if (d >= 0 && d < 3) k.dimvar[t_[i].s] = d;
      }
    }
    for (size_t i = f.lbrace + 1; i < f.rbrace; ++i)
      if (t_[i].s == "if" && ClampIf(k, i, f.rbrace) >= 0) ++k.clamp_ifs;
  }
}

// Recognises a clamped-face read of the staged grid along axis d:
//   if (i == 0)     v = PSGridGet(g, i, j, k); else v = PSGridGet(g, i-1, j, k);
//   if (i == n - 1) v = PSGridGet(g, i, j, k); else v = PSGridGet(g, i+1, j, k);
// (n a local holding PSGridDim(g, d), or PSGridDim(g, d) itself), i.e.
// v = g at the clamped coordinate i-1 / i+1.  Returns d, or -1.
int Translator::ClampIf(const Kernel &k, size_t i, size_t e) const {
  if (k.staged.empty() || k.rank != 3 || i + 1 >= e || t_[i].s != "if" || t_[i + 1].s != "(") return -1;
  const size_t cl = MatchClose(t_, i + 1);
  if (cl >= e) return -1;
  size_t c = i + 2;
  int d = -1;
  for (int a = 0; a < 3; ++a)
    if (t_[c].kind == Tok::kIdent && t_[c].s == k.idx[a]) d = a;
  if (d < 0 || t_[c + 1].s != "==") return -1;
  c += 2;
  bool low;
  if (cl - c == 1 && t_[c].kind == Tok::kNum && t_[c].s == "0") {
    low = true;
  } else if (cl - c == 3 && t_[c].kind == Tok::kIdent && k.dimvar.count(t_[c].s) && k.dimvar.at(t_[c].s) == d &&
             t_[c + 1].s == "-" && t_[c + 2].s == "1") {
    low = false;
  } else if (cl - c == 8 && t_[c].s == "PSGridDim" && t_[c + 1].s == "(" && t_[c + 2].s == k.staged &&
             t_[c + 3].s == "," && t_[c + 4].kind == Tok::kNum && atoi(t_[c + 4].s.c_str()) == d &&
             t_[c + 5].s == ")" && t_[c + 6].s == "-" && t_[c + 7].s == "1") {
    low = false;
  } else {
    return -1;
  }
  // "v = PSGridGet(g, ...);" with the given offsets; returns the index after
  // the ';' or 0.
  auto assign = [&](size_t q, const std::string &v, const int (&want)[3], std::string *name) -> size_t {
    if (q + 3 >= e || t_[q].kind != Tok::kIdent || t_[q + 1].s != "=" || t_[q + 2].s != "PSGridGet" ||
        t_[q + 3].s != "(")
      return 0;
    if (!v.empty() && t_[q].s != v) return 0;
    *name = t_[q].s;
    const size_t close = MatchClose(t_, q + 3);
    if (close + 1 >= e || t_[close + 1].s != ";") return 0;
    auto args = SplitArgs(t_, q + 4, close);
    if (args.size() != 4 || Text(t_, args[0].first, args[0].second) != k.staged) return 0;
    for (int a = 0; a < 3; ++a) {
      int off;
      if (!RelativeIndex(t_, args[a + 1].first, args[a + 1].second, k.idx[a], &off) || off != want[a]) return 0;
    }
    return close + 2;
  };
  const int zero[3] = {0, 0, 0};
  int nb[3] = {0, 0, 0};
  nb[d] = low ? -1 : 1;
  std::string v;
  size_t q = assign(cl + 1, "", zero, &v);
  if (!q || q >= e || t_[q].s != "else") return -1;
  std::string v2;
  if (!assign(q + 1, v, nb, &v2)) return -1;
  return d;
}

bool Translator::EmitIsTail(const Func &f) const {
  size_t emit_at = 0;
  int depth = 0;
  bool top = false;
  for (size_t i = f.lbrace + 1; i < f.rbrace; ++i) {
    const std::string &s = t_[i].s;
    if (s == "{") ++depth;
    else if (s == "}") --depth;
    else if (s == "PSGridEmit") {
      emit_at = i;
      top = depth == 0;
    } else if (s == "return" && i + 1 < f.rbrace && t_[i + 1].s != ";") {
      top = false;  // value returns are not part of the DSL
    }
  }
  if (!top || !emit_at) return false;
  size_t close = MatchClose(t_, emit_at + 1);
  size_t q = close + 1;
  if (q < f.rbrace && t_[q].s == ";") ++q;
  if (q < f.rbrace && t_[q].s == "return" && q + 1 < f.rbrace && t_[q + 1].s == ";") q += 2;
  if (q != f.rbrace) return false;
  for (size_t i = f.lbrace + 1; i < emit_at; ++i)  // no early "return;"
    if (t_[i].s == "return") return false;
  return true;
}

std::string Translator::RewriteBody(const Kernel &k, size_t b, size_t e, bool compute_form, bool exact) {
  std::string out;
  int pdepth = 0;
  auto add = [&](const std::string &s) {
    if (!out.empty() && IsIdentChar(out.back()) && !s.empty() && IsIdentChar(s[0])) out += ' ';
    out += s;
  };
  const std::string uref = UniformRef(k);
  // Offset of the map's current point in grid g.
  auto center = [&](const std::string &g, const GridType &gt) {
    std::string at = "psk::Off" + std::to_string(gt.rank) + "(" + g;
    for (int d = 0; d < gt.rank; ++d) at += ", " + k.idx[std::min(d, k.rank - 1)];
    return at + ")";
  };
  // Member access `.name[s0][s1]..` of a point-type grid at tokens [q, e):
  // returns the element's address expression and advances q past it.
  auto member = [&](const std::string &g, const UType &u, size_t &q, const std::string &off) {
    if (q + 1 >= e || t_[q].s != "." || t_[q + 1].kind != Tok::kIdent)
      Die("kernel " + k.f->name + ": expected a member of point type " + u.ctype + " after '" + g + "'");
    const UMember *m = u.Find(t_[q + 1].s);
    if (!m) Die("kernel " + k.f->name + ": " + u.ctype + " has no member '" + t_[q + 1].s + "'");
    q += 2;
    std::vector<std::string> subs;
    while (q < e && t_[q].s == "[") {
      size_t c = MatchClose(t_, q);
      subs.push_back(RewriteBody(k, q + 1, c, false));
      q = c + 1;
    }
    if (subs.size() != m->dims.size())
      Die("kernel " + k.f->name + ": member '" + m->name + "' must be indexed down to a scalar");
    std::string idx = subs.empty() ? off : "(int64_t)" + MemberIndex(*m, subs) + " * " + g + ".ps + " + off;
    return std::make_pair(m, g + "." + m->name + "[" + idx + "]");
  };
  for (size_t i = b; i < e; ++i) {
    const Tok &tk = t_[i];
    if (exact && tk.kind == Tok::kIdent && tk.s == "if") {
      const int d = ClampIf(k, i, e);
      if (d >= 0) {
        // The skeleton supplies the clamped-coordinate neighbour along axis
        // d when __kExact<d> is set, so the condition folds to false there.
        add("if");
        add("(");
        add(std::string("!__kExact") + "XYZ"[d]);
        add("&&");
        ++pdepth;  // stands for the condition's own "(" skipped here
        ++i;
        continue;
      }
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && tk.s == "PSGridEmitUtype") {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      if (args.size() != 2) Die("kernel " + k.f->name + ": PSGridEmitUtype takes a member and a value");
      std::string g = t_[args[0].first].s;
      const GridType &gt = k.gtype.at(g);
      const UType &u = g_utypes.at("PSGrid" + std::to_string(gt.rank) + "D" + gt.utype);
      size_t q = args[0].first + 1;
      auto mem = member(g, u, q, center(g, gt));
      if (q != args[0].second) Die("kernel " + k.f->name + ": unexpected tokens in the PSGridEmitUtype target");
      add("(" + mem.second + " = (" + mem.first->ctype + ")(" +
          RewriteBody(k, args[1].first, args[1].second, false) + "))");
      i = close;
      continue;
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" &&
        (tk.s == "PSGridGet" || tk.s == "PSGridGetPeriodic" || tk.s == "PSGridEmit" || tk.s == "PSGridDim")) {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      std::string g = Text(t_, args[0].first, args[0].second);
      const GridType &gt = k.gtype.at(g);
      std::vector<std::string> a;
      for (size_t q = 1; q < args.size(); ++q) a.push_back(RewriteBody(k, args[q].first, args[q].second, false));
      std::ostringstream r;
      if (!gt.utype.empty() && tk.s != "PSGridDim") {
        const UType &u = g_utypes.at("PSGrid" + std::to_string(gt.rank) + "D" + gt.utype);
        if (tk.s == "PSGridEmit") {
          r << "__PSUPut_" << u.name << "(" << g << ", " << center(g, gt) << ", " << a[0] << ")";
        } else {
          std::string off = "psk::Off" + std::string(tk.s == "PSGridGetPeriodic" ? "P" : "") +
                            std::to_string(gt.rank) + "(" + g;
          for (auto &s : a) off += ", " + s;
          off += ")";
          size_t q = close + 1;
          if (q < e && t_[q].s == ".") {
            r << "(" << member(g, u, q, off).second << ")";
            close = q - 1;
          } else {
            r << "__PSUGet_" << u.name << "(" << g << ", " << off << ")";
          }
        }
        add(r.str());
        i = close;
        continue;
      }
      if (tk.s == "PSGridDim") {
        r << "(" << g << ".dim[" << a[0] << "])";
      } else if (tk.s == "PSGridEmit") {
        const std::string at = center(g, gt);
        if (compute_form) {
          r << "return (" << gt.elem << ")(" << a[0] << ")";
        } else if (!uref.empty() && gt.rank == 3) {
          r << "psk::Put3<__kU>(" << g << ", __ref, " << k.idx[0] << ", " << k.idx[1] << ", " << k.idx[2] << ", ("
            << gt.elem << ")(" << a[0] << "))";
        } else {
          r << g << ".store(" << at << ", (" << gt.elem << ")(" << a[0] << "))";
        }
      } else {
        Get acc;
        acc.relative = gt.rank <= k.rank;
        for (int d = 0; d < gt.rank && acc.relative; ++d)
          acc.relative = RelativeIndex(t_, args[d + 1].first, args[d + 1].second, k.idx[d], &acc.off[d]);
        if (point_mode_) {
          if (g == k.staged) {
            r << "((float)__PSat(" << acc.off[0] << ", " << acc.off[1] << ", " << acc.off[2] << "))";
          } else {
            const auto it = std::find(k.point_grids.begin(), k.point_grids.end(), g);
            r << "((float)__PSpt(" << (it - k.point_grids.begin()) << "))";
          }
        } else if (g == k.staged && acc.relative) {
          r << "((" << gt.elem << ")__ctx.at(0, " << acc.off[0] << ", " << acc.off[1] << ", " << acc.off[2]
            << "))";
        } else if (!uref.empty() && gt.rank == 3 && tk.s != "PSGridGetPeriodic") {
          r << "psk::Get3<__kU>(" << g << ", __ref";
          for (auto &s : a) r << ", " << s;
          r << ")";
        } else {
          r << g << ".load(psk::Off" << (tk.s == "PSGridGetPeriodic" ? "P" : "") << gt.rank << "(" << g;
          for (auto &s : a) r << ", " << s;
          r << "))";
        }
      }
      add(r.str());
      i = close;
      continue;
    }
    if (compute_form && tk.kind == Tok::kIdent && tk.s == "return" && i + 1 < e && t_[i + 1].s == ";") {
      ++i;  // the trailing "return;" of a star kernel
      continue;
    }
    if (tk.kind == Tok::kPunct && tk.s == "(") ++pdepth;
    if (tk.kind == Tok::kPunct && tk.s == ")") --pdepth;
    if (tk.kind == Tok::kPunct && ((tk.s == ";" && pdepth == 0) || tk.s == "{" || tk.s == "}")) {
      add(tk.s);
      out += "\n";
      continue;
    }
    if (!out.empty() && out.back() != '\n' && out.back() != '(' && tk.s != ")" && tk.s != "," && tk.s != ";")
      out += ' ';
    add(tk.s);
  }
  return out;
}

std::string Translator::Prologue(const Kernel &k) const {
  std::ostringstream q;
  for (const Param &p : k.all) {
    if (p.is_grid) q << "    const auto &" << p.name << " = __a." << p.name << ";\n";
    else q << "    " << p.type << " " << p.name << " = __a." << p.name << ";\n";
  }
  for (const Param &p : k.all) q << "    (void)" << p.name << ";\n";
  const std::string uref = UniformRef(k);
  if (!uref.empty()) q << "    const auto &__ref = __a." << uref << ";\n    (void)__ref;\n";
  return q.str();
}

std::string Translator::EmitPlan(const Kernel &k) {
  const Func &f = *k.f;
  std::ostringstream o;
  const std::string P = "__PSPlan_" + f.name;
  std::string sv = k.staged.empty() ? "float" : k.gtype.at(k.staged).elem;
  o << "// Plan for kernel " << f.name << " (rank " << k.rank << ", staged grid: "
    << (k.staged.empty() ? "none" : k.staged) << ", radius " << k.rx << "/" << k.ry << "/" << k.rz
    << (k.star ? ", star form" : "") << (k.fusable ? ", fusable pairs" : "") << ")\n";
  // __kU: launched with uniform extents (every rank-3 grid has the extents
  // of __ref and fits 32-bit byte offsets): accesses at a point share one
  // element index (psk::Get3/Put3).
  o << "template <class SF, bool __kU = false>\nstruct " << P << " {\n";
  o << "  using SV = " << sv << ";\n";
  o << "  static constexpr int NS = " << (k.staged.empty() ? 0 : 1) << ", RX = " << k.rx << ", RY = " << k.ry
    << ", RZ = " << k.rz << ";\n";
  o << "  static constexpr bool kWrap = " << (k.wrap ? "true" : "false") << ";\n";
  o << "  static constexpr bool kReduce = false;\n";
  // Grids the body touches straight in HBM (not staged); a user-type grid
  // counts as several streams.  The launcher gives a thread two tile rows
  // only when this is small (each row's direct loads take registers).
  int direct = 0;
  for (const Param &p : k.all)
    if (p.is_grid && p.name != k.staged) direct += p.gt.utype.empty() ? 1 : 4;
  o << "  static constexpr int kDirectGrids = " << direct << ";\n";
  o << "  struct Args {\n";
  for (const Param &p : k.all) {
    if (p.is_grid && !p.gt.utype.empty()) o << "    struct __PSURef_" << p.gt.utype << " " << p.name << ";\n";
    else if (p.is_grid) o << "    psk::GridRef<" << StorageOf(p.gt) << "> " << p.name << ";\n";
    else o << "    " << p.type << " " << p.name << ";\n";
  }
  o << "  };\n";
  if (!k.staged.empty()) {
    const GridType &sg = k.gtype.at(k.staged);
    o << "  using Storage = " << StorageOf(sg) << ";\n";
    o << "  __device__ static void stage_dims(const Args &a, int &nx, int &ny, int &nz) {\n"
      << "    nx = a." << k.staged << ".dim[0]; ny = a." << k.staged << ".dim[1]; nz = a." << k.staged
      << ".dim[2];\n  }\n";
    o << "  __device__ static SV stage_load(const Args &a, int, int64_t off) { return a." << k.staged
      << ".load(off); }\n";
    o << "  __host__ __device__ static const typename Storage::stored_t *stage_ptr(const Args &a) { return a."
      << k.staged << ".p; }\n";
    o << "  static void set_in(Args &a, const typename Storage::stored_t *p) { a." << k.staged
      << ".p = const_cast<typename Storage::stored_t *>(p); }\n";
  }
  if (k.star) {
    o << "  __host__ __device__ static typename Storage::stored_t *out_ptr(const Args &a) { return a."
      << k.out_grid << ".p; }\n";
    o << "  static void set_out(Args &a, typename Storage::stored_t *p) { a." << k.out_grid << ".p = p; }\n";
  }
  auto prologue = [&]() { return Prologue(k); };
  std::string params = "const C &__ctx, const Args &__a";
  const char *dummies[3] = {"__i0", "__i1", "__i2"};
  for (int d = 0; d < 3; ++d) params += ", const int " + (d < k.rank ? k.idx[d] : std::string(dummies[d]));
  o << "  template <class C>\n  __device__ static void point(" << params << ") {\n";
  o << "    (void)__ctx;\n" << prologue();
  o << "    {\n" << RewriteBody(k, f.lbrace + 1, f.rbrace, false) << "    }\n  }\n";
  if (k.star) {
    o << "  template <class C>\n  __device__ static float compute(" << params << ") {\n";
    o << "    (void)__ctx;\n" << prologue();
    o << "    {\n" << RewriteBody(k, f.lbrace + 1, f.rbrace, true) << "    }\n  }\n";
  }
  if (k.star && k.clamp_ifs > 0) {
    // compute() for a point whose neighbours along the flagged axes are the
    // staged grid's elements at clamped coordinates (skeleton_star_tb.h):
    // the clamped-face reads then take the neighbour directly.
    o << "  static constexpr bool kExactNeighbours = true;\n";
    o << "  template <bool __kExactX, bool __kExactY, bool __kExactZ, class C>\n  __device__ static float "
         "compute_exact("
      << params << ") {\n";
    o << "    (void)__ctx;\n" << prologue();
    o << "    {\n" << RewriteBody(k, f.lbrace + 1, f.rbrace, true, true) << "    }\n  }\n";
  }
  if (k.point) {
    // psk::PointTB's plan interface: the point grids in parameter order, the
    // scalars, and the body as update(): __PSpt(n) is point grid n at the
    // point, __PSat(dx, dy, dz) the staged grid at an offset; the body is
    // evaluated as written whatever the storage (as point() evaluates it).
    o << "  static constexpr int NG = " << k.point_grids.size() << ";\n";
    o << "  struct Scalars {\n";
    int nsc = 0;
    for (const Param &p : k.all)
      if (!p.is_grid) {
        o << "    " << StripConst(p.type) << " " << p.name << ";\n";
        ++nsc;
      }
    if (!nsc) o << "    int __none;\n";
    o << "  };\n";
    o << "  template <bool __kFma, class __G, class __PF>\n"
      << "  __device__ static float update(const __G &__PSpt, const __PF &__PSat, const Scalars &__s, float *) {\n";
    for (const Param &p : k.all)
      if (!p.is_grid) o << "    " << p.type << " " << p.name << " = __s." << p.name << ";\n    (void)" << p.name << ";\n";
    o << "    (void)__s;\n";
    point_mode_ = true;
    o << "    {\n" << RewriteBody(k, f.lbrace + 1, f.rbrace, true) << "    }\n  }\n";
    point_mode_ = false;
  }
  o << "};\n\n";
  return o.str();
}

std::string Translator::EmitReduceKernel(const Kernel &k) {
  const Func &f = *k.f;
  const std::string T = t_[f.lparen - 2].s;
  const std::string P = "__PSRPlan_" + f.name;
  std::ostringstream o;
  o << "// Reduction kernel " << f.name << " (rank " << k.rank << ", value " << T << ")\n";
  o << "template <class SF, bool __kU = false>\nstruct " << P << " {\n";
  o << "  struct Args {\n    using SFT = SF;\n";
  for (const Param &p : k.all) {
    if (p.is_grid && !p.gt.utype.empty()) o << "    struct __PSURef_" << p.gt.utype << " " << p.name << ";\n";
    else if (p.is_grid) o << "    psk::GridRef<" << StorageOf(p.gt) << "> " << p.name << ";\n";
    else o << "    " << StripConst(p.type) << " " << p.name << ";\n";
  }
  o << "  };\n";
  std::string params = "const Args &__a";
  const char *dummies[3] = {"__i0", "__i1", "__i2"};
  for (int d = 0; d < 3; ++d) params += ", const int " + (d < k.rank ? k.idx[d] : std::string(dummies[d]));
  o << "  __device__ static " << T << " value(" << params << ") {\n" << Prologue(k);
  o << "    {\n" << RewriteBody(k, f.lbrace + 1, f.rbrace, false) << "    }\n  }\n};\n\n";
  std::string decl, names;
  for (const Param &p : k.all) {
    decl += std::string(", ") + (p.is_grid ? "__PSGrid *" : p.type + " ") + p.name;
    names += ", " + p.name;
  }
  // The plan's Args for either instantiation (__kU or not).
  o << "template <class __A>\nstatic void __PSRFill_" << f.name << "(__A &__a" << decl << ") {\n";
  o << "  using SF = typename __A::SFT;\n";
  for (const Param &p : k.all) {
    if (p.is_grid && !p.gt.utype.empty())
      o << "  __a." << p.name << " = __PSUMakeRef_" << p.gt.utype << "(" << p.name << ");\n";
    else if (p.is_grid)
      o << "  __a." << p.name << " = psk::MakeRef<" << (p.gt.elem == "float" ? "SF" : StorageOf(p.gt)) << ">("
        << p.name << ");\n";
    else
      o << "  __a." << p.name << " = " << p.name << ";\n";
  }
  o << "  (void)sizeof(SF);\n}\n\n";
  o << "template <class SF>\nstatic void __PSReduceMapT_" << f.name << "(" << T
    << " *__out, enum PSReduceOp __op, __PSDomain __dom" << decl << ") {\n";
  o << "  const psk::Box __box = psk::BoxFromDomain(__dom, " << k.rank << ");\n";
  o << "  double __r = 0.0;\n  hipError_t __e;\n";
  o << "  typename " << P << "<SF>::Args __a;\n  __PSRFill_" << f.name << "(__a" << names << ");\n";
  const std::string uref = UniformRef(k);
  if (!uref.empty()) {
    // Every rank-3 grid with the extents of one of them and 32-bit byte
    // offsets: one element index per point (psk::Get3 with __kU).
    o << "  if (psk::UniformExtents(__a." << uref;
    for (const Param &p : k.all)
      if (p.is_grid && p.gt.rank == 3 && p.gt.utype.empty() && p.name != uref) o << ", __a." << p.name;
    o << ")) {\n    typename " << P << "<SF, true>::Args __u;\n    __PSRFill_" << f.name << "(__u" << names
      << ");\n    __e = psk::RunReduceMap<" << P << "<SF, true>>(__u, __box, __op, &__r, (hipStream_t)__PSGetStream());\n"
      << "  } else {\n  ";
  }
  o << "  __e = psk::RunReduceMap<" << P << "<SF>>(__a, __box, __op, &__r, (hipStream_t)__PSGetStream());\n";
  if (!uref.empty()) o << "  }\n";
  o << "  psk::CheckLaunch(__e, \"" << f.name << "\");\n";
  o << "  *__out = (" << T << ")__r;\n}\n\n";
  std::vector<std::string> floats;
  for (const Param &p : k.all)
    if (p.is_grid && p.gt.elem == "float") floats.push_back(p.name);
  o << "static void __PSReduceMap_" << f.name << "(" << T << " *__out, enum PSReduceOp __op, __PSDomain __dom"
    << decl << ") {\n";
  o << "  if (__op < PS_MAX || __op > PS_PROD) PSAbort(1);\n";
  if (floats.empty()) {
    o << "  __PSReduceMapT_" << f.name << "<psk::Native<float>>(__out, __op, __dom" << names << ");\n";
  } else {
    o << "  const __PSGrid *__g[] = {";
    for (size_t i = 0; i < floats.size(); ++i) o << (i ? ", " : "") << floats[i];
    o << "};\n  switch (psk::FloatStorage(__g, " << floats.size() << ")) {\n";
    o << "    case PS_HIP_STORAGE_BF16: __PSReduceMapT_" << f.name << "<psk::BF16>(__out, __op, __dom" << names
      << "); break;\n";
    o << "    case PS_HIP_STORAGE_FP16: __PSReduceMapT_" << f.name << "<psk::FP16>(__out, __op, __dom" << names
      << "); break;\n";
    o << "    default: __PSReduceMapT_" << f.name << "<psk::Native<float>>(__out, __op, __dom" << names
      << "); break;\n  }\n";
  }
  if (cfg_.trace)
    o << "  if (__ps_trace) fprintf(__ps_trace, \"Physis: Reduce " << f.name << " finished\\n\");\n";
  o << "}\n\n";
  return o.str();
}

std::string Translator::EmitMapAndLaunch(const Kernel &k) {
  const Func &f = *k.f;
  std::ostringstream o;
  const std::string S = "__PSStencil_" + f.name, P = "__PSPlan_" + f.name;
  // The stencil object (reference_runtime_builder.cc:398-448): domain, each
  // grid with its id, then the scalars.
  o << "struct " << S << " {\n  __PSDomain dom;\n";
  for (const Param &p : k.all) {
    if (p.is_grid) o << "  __PSGrid *" << p.name << ";\n  int " << p.name << "_index;\n";
    else o << "  " << p.type << " " << p.name << ";\n";
  }
  o << "};\n\n";
  o << "static struct " << S << " __PSStencilMap_" << f.name << "(__PSDomain dom";
  for (const Param &p : k.all) o << ", " << (p.is_grid ? "__PSGrid *" : p.type + " ") << p.name;
  o << ") {\n  struct " << S << " s;\n  s.dom = dom;\n";
  for (const Param &p : k.all) {
    if (p.is_grid) o << "  s." << p.name << " = " << p.name << ";\n  s." << p.name << "_index = __PSGridGetID("
                     << p.name << ");\n";
    else o << "  s." << p.name << " = " << p.name << ";\n";
  }
  o << "  return s;\n}\n\n";
  // Typed args.
  o << "template <class SF, bool __kU = false>\nstatic typename " << P << "<SF, __kU>::Args __PSArgs_" << f.name
    << "(const struct " << S << " *s) {\n  typename " << P << "<SF, __kU>::Args a;\n";
  for (const Param &p : k.all) {
    if (p.is_grid && !p.gt.utype.empty())
      o << "  a." << p.name << " = __PSUMakeRef_" << p.gt.utype << "(s->" << p.name << ");\n";
    else if (p.is_grid) o << "  a." << p.name << " = psk::MakeRef<" << (p.gt.elem == "float" ? "SF" : StorageOf(p.gt))
                     << ">(s->" << p.name << ");\n";
    else o << "  a." << p.name << " = s->" << p.name << ";\n";
  }
  o << "  return a;\n}\n\n";
  // One map application.
  // parity: -1 all points; 0/1 the red/black points of a red-black map.
  o << "template <class SF>\nstatic void __PSLaunchT_" << f.name << "(const struct " << S
    << " *s, hipStream_t st, int parity) {\n";
  o << "  typename " << P << "<SF>::Args a = __PSArgs_" << f.name << "<SF>(s);\n";
  o << "  psk::Box box = psk::BoxFromDomain(s->dom, " << k.rank << ");\n";
  if (k.star)
    o << "  if (parity < 0) {\n    psk::CheckLaunch(psk::LaunchStarStep<" << P << "<SF>>(a, box, a." << k.staged
      << ".dim[0], a." << k.staged << ".dim[1], a." << k.staged << ".dim[2], st), \"" << f.name
      << "\");\n    return;\n  }\n";
  const std::string uref = UniformRef(k);
  if (!uref.empty()) {
    o << "  if (psk::UniformExtents(a." << uref;
    for (const Param &p : k.all)
      if (p.is_grid && p.gt.rank == 3 && p.gt.utype.empty() && p.name != uref) o << ", a." << p.name;
    o << ")) {\n    psk::CheckLaunch(psk::LaunchRing<" << P << "<SF, true>, 3>(__PSArgs_" << f.name
      << "<SF, true>(s), box, st, parity), \"" << f.name << "\");\n    return;\n  }\n";
  }
  o << "  psk::CheckLaunch(psk::LaunchRing<" << P << "<SF>, " << k.rank << ">(a, box, st, parity), \"" << f.name
    << "\");\n";
  o << "}\n\n";
  // Storage dispatch.
  std::vector<std::string> floats;
  for (const Param &p : k.all)
    if (p.is_grid && p.gt.elem == "float") floats.push_back(p.name);
  o << "static int __PSStorage_" << f.name << "(const struct " << S << " *s) {\n";
  if (floats.empty()) {
    o << "  (void)s;\n  return PS_HIP_STORAGE_NATIVE;\n";
  } else {
    o << "  const __PSGrid *g[] = {";
    for (size_t i = 0; i < floats.size(); ++i) o << (i ? ", " : "") << "s->" << floats[i];
    o << "};\n  return psk::FloatStorage(g, " << floats.size() << ");\n";
  }
  o << "}\n\n";
  o << "static void __PSLaunch_" << f.name << "(const struct " << S << " *s, int parity) {\n";
  o << "  hipStream_t st = (hipStream_t)__PSGetStream();\n";
  o << "  switch (__PSStorage_" << f.name << "(s)) {\n";
  o << "    case PS_HIP_STORAGE_BF16: __PSLaunchT_" << f.name << "<psk::BF16>(s, st, parity); break;\n";
  o << "    case PS_HIP_STORAGE_FP16: __PSLaunchT_" << f.name << "<psk::FP16>(s, st, parity); break;\n";
  o << "    default: __PSLaunchT_" << f.name << "<psk::Native<float>>(s, st, parity); break;\n";
  o << "  }\n";
  if (cfg_.error_check) o << "  __PSCheckDeviceError(\"" << f.name << "\");\n";
  o << "}\n\n";
  if (k.fusable) {
    // Ping-pong fusion: two maps of this kernel that differ only by swapping
    // the staged and the emitted grid, over the whole grid.  Other grids the
    // body reads must be neither of the two (a fused pass reads them while
    // the swapped grids hold other time levels than separate steps see).
    o << "static int __PSFusable_" << f.name << "(const struct " << S << " *a, const struct " << S << " *b) {\n";
    o << "  if (a->" << k.staged << " != b->" << k.out_grid << " || a->" << k.out_grid << " != b->" << k.staged
      << " || a->" << k.staged << " == a->" << k.out_grid << ") return 0;\n";
    for (const Param &p : k.all) {
      if (p.name == k.staged || p.name == k.out_grid) continue;
      if (p.is_grid) {
        o << "  if (a->" << p.name << " != b->" << p.name << ") return 0;\n";
        o << "  if (a->" << p.name << " == a->" << k.staged << " || a->" << p.name << " == a->" << k.out_grid
          << ") return 0;\n";
      } else {
        o << "  if (memcmp(&a->" << p.name << ", &b->" << p.name << ", sizeof(a->" << p.name
          << ")) != 0) return 0;\n";
      }
    }
    o << "  const __PSGrid *g = a->" << k.staged << ";\n";
    o << "  for (int d = 0; d < 3; ++d) {\n"
      << "    if (a->dom.local_min[d] != 0 || a->dom.local_max[d] != g->dim[d]) return 0;\n"
      << "    if (b->dom.local_min[d] != 0 || b->dom.local_max[d] != g->dim[d]) return 0;\n  }\n";
    o << "  return psk::FusedPairSupported(g->dim[0], g->dim[1], g->dim[2]);\n}\n\n";
    o << "template <class SF>\nstatic void __PSRunPingPongT_" << f.name << "(const struct " << S
      << " *a, const struct " << S << " *b, int iter, hipStream_t st) {\n";
    o << "  const __PSGrid *g = a->" << k.staged << ";\n";
    o << "  size_t bytes = (size_t)g->num_elms * PSHipGridStoredElmSize(g);\n";
    o << "  psk::RunPingPong<" << P << "<SF>>(__PSArgs_" << f.name << "<SF>(a), __PSArgs_" << f.name
      << "<SF>(b), iter, g->dim[0], g->dim[1], g->dim[2], bytes, st);\n}\n\n";
    o << "static void __PSRunPingPong_" << f.name << "(const struct " << S << " *a, const struct " << S
      << " *b, int iter) {\n";
    o << "  hipStream_t st = (hipStream_t)__PSGetStream();\n";
    o << "  switch (__PSStorage_" << f.name << "(a)) {\n";
    o << "    case PS_HIP_STORAGE_BF16: __PSRunPingPongT_" << f.name << "<psk::BF16>(a, b, iter, st); break;\n";
    o << "    case PS_HIP_STORAGE_FP16: __PSRunPingPongT_" << f.name << "<psk::FP16>(a, b, iter, st); break;\n";
    o << "    default: __PSRunPingPongT_" << f.name << "<psk::Native<float>>(a, b, iter, st); break;\n";
    o << "  }\n";
    if (cfg_.error_check) o << "  __PSCheckDeviceError(\"" << f.name << "\");\n";
    o << "}\n\n";
  }
  if (k.point) {
    // Two maps of a point-form kernel that differ only by swapping the staged
    // and the emitted grid, over the interior of the staged grid, FP32
    // storage, every grid of one shape below 4 GiB: two sweeps per HBM pass
    // (psk::RunPointTB), leaving both grids as the separate sweeps do.
    const std::string &in = k.staged, &out = k.out_grid;
    o << "static int __PSPointFusable_" << f.name << "(const struct " << S << " *a, const struct " << S
      << " *b) {\n";
    o << "  if (a->" << in << " != b->" << out << " || a->" << out << " != b->" << in << " || a->" << in
      << " == a->" << out << ") return 0;\n";
    for (const Param &p : k.all) {
      if (p.name == in || p.name == out) continue;
      if (p.is_grid) {
        o << "  if (a->" << p.name << " != b->" << p.name << ") return 0;\n";
        o << "  if (a->" << p.name << " == a->" << in << " || a->" << p.name << " == a->" << out << ") return 0;\n";
      } else {
        o << "  if (memcmp(&a->" << p.name << ", &b->" << p.name << ", sizeof(a->" << p.name << ")) != 0) return 0;\n";
      }
    }
    o << "  if (__PSStorage_" << f.name << "(a) != PS_HIP_STORAGE_NATIVE) return 0;\n";
    o << "  const __PSGrid *g = a->" << in << ";\n";
    o << "  for (int d = 0; d < 3; ++d) {\n"
      << "    if (a->dom.local_min[d] != 1 || a->dom.local_max[d] != g->dim[d] - 1) return 0;\n"
      << "    if (b->dom.local_min[d] != 1 || b->dom.local_max[d] != g->dim[d] - 1) return 0;\n  }\n";
    o << "  const __PSGrid *gs[] = {";
    bool first = true;
    for (const Param &p : k.all)
      if (p.is_grid) {
        o << (first ? "" : ", ") << "a->" << p.name;
        first = false;
      }
    o << "};\n";
    o << "  for (const __PSGrid *x : gs) {\n"
      << "    if (x->num_dims != 3 || x->dim[0] != g->dim[0] || x->dim[1] != g->dim[1] || x->dim[2] != g->dim[2]) "
         "return 0;\n"
      << "    if ((uint64_t)x->num_elms * sizeof(float) > 0xffffffffull) return 0;\n  }\n";
    o << "  return 1;\n}\n\n";
    o << "static void __PSRunPointTB_" << f.name << "(const struct " << S << " *a, int iter) {\n";
    o << "  using SF = psk::Native<float>;\n";
    o << "  psk::PointTBArgs<typename " << P << "<SF>::Scalars> t{};\n";
    o << "  const __PSGrid *g = a->" << in << ";\n";
    o << "  for (int d = 0; d < 3; ++d) t.dim[d] = g->dim[d];\n";
    for (size_t n = 0; n < k.point_grids.size(); ++n)
      o << "  t.c[" << n << "] = a->" << k.point_grids[n] << "->p0;\n";
    for (const Param &p : k.all)
      if (!p.is_grid) o << "  t.sc." << p.name << " = a->" << p.name << ";\n";
    o << "  void *X = __PSHipAuxBuffer((size_t)g->num_elms * sizeof(float));\n";
    o << "  if (!X) {\n    fprintf(stderr, \"physis-hip: no memory for the fused-sweep buffer\\n\");\n"
      << "    PSAbort(1);\n  }\n";
    o << "  psk::CheckLaunch(psk::RunPointTB<" << P << "<SF>, SF>(t, a->" << in << "->p0, a->" << out
      << "->p0, X, 2 * iter, (hipStream_t)__PSGetStream()), \"" << f.name << "\");\n";
    o << "  if (__ps_trace) fprintf(__ps_trace, \"Physis: fused sweeps: %d in passes of two on the point tile\\n\", "
         "2 * iter);\n";
    if (cfg_.error_check) o << "  __PSCheckDeviceError(\"" << f.name << "\");\n";
    o << "}\n\n";
  }
  return o.str();
}

std::string Translator::EmitRun(const std::vector<std::pair<std::string, std::string>> &maps,
                                const std::vector<int> &modes, int id) {
  // maps: (kernel name, argument text of the map call)
  std::ostringstream o;
  o << "static float __PSStencilRun_" << id << "(int iter";
  for (size_t i = 0; i < maps.size(); ++i) o << ", struct __PSStencil_" << maps[i].first << " s" << i;
  o << ") {\n";
  if (cfg_.trace) {
    o << "  __PSTraceStencilPre(\"";
    for (auto &m : maps) o << m.first;
    o << "\");\n  __PSStopwatch __st;\n  __PSStopwatchStart(&__st);\n";
  }
  bool fusable = cfg_.fuse && maps.size() == 2 && maps[0].first == maps[1].first &&
                 kernels_.at(maps[0].first).fusable && modes[0] == 0 && modes[1] == 0;
  bool point = cfg_.fuse && maps.size() == 2 && maps[0].first == maps[1].first && kernels_.at(maps[0].first).point &&
               modes[0] == 0 && modes[1] == 0;
  if (fusable) {
    const std::string &kn = maps[0].first;
    o << "  if (iter > 1 && __PSFusable_" << kn << "(&s0, &s1)) {\n";
    o << "    __PSRunPingPong_" << kn << "(&s0, &s1, iter);\n  } else {\n";
  } else if (point) {
    const std::string &kn = maps[0].first;
    o << "  if (iter > 1 && __PSPointFusable_" << kn << "(&s0, &s1)) {\n";
    o << "    __PSRunPointTB_" << kn << "(&s0, iter);\n  } else {\n";
  }
  o << "  for (int i = 0; i < iter; ++i) {\n";
  // A red-black map runs its red then its black pass (reference
  // reference_translator.cc:628-640); MapRed / MapBlack run one colour.
  for (size_t i = 0; i < maps.size(); ++i) {
    const std::string call = "    __PSLaunch_" + maps[i].first + "(&s" + std::to_string(i);
    switch (modes[i]) {
      case 1: o << call << ", 0);\n" << call << ", 1);\n"; break;
      case 2: o << call << ", 0);\n"; break;
      case 3: o << call << ", 1);\n"; break;
      default: o << call << ", -1);\n"; break;
    }
  }
  o << "  }\n";
  if (fusable || point) o << "  }\n";
  if (cfg_.trace) {
    o << "  (void)hipStreamSynchronize((hipStream_t)__PSGetStream());\n";
    o << "  float f = __PSStopwatchStop(&__st);\n  __PSTraceStencilPost(f);\n  return f;\n";
  } else {
    o << "  return 0.0f;\n";
  }
  o << "}\n\n";
  (void)id;
  return o.str();
}

std::string Translator::RewriteHost(size_t b, size_t e) {
  std::string out;
  int pdepth = 0;
  auto add = [&](const std::string &s) {
    if (!out.empty()) {
      char p = out.back();
      if ((IsIdentChar(p) && IsIdentChar(s[0])) || (p != '\n' && p != '(' && p != '[' && s != ")" && s != "," &&
                                                    s != ";" && s != "]" && s != "(" && s != "[" && s != "."))
        out += ' ';
    }
    out += s;
  };
  for (size_t i = b; i < e; ++i) {
    const Tok &tk = t_[i];
    if (tk.kind == Tok::kPunct && tk.s.size() > 1 && tk.s[0] == '#') {
      if (!out.empty() && out.back() != '\n') out += "\n";
      out += tk.s;
      continue;
    }
    GridType gt;
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && tk.s.size() == 13 &&
        tk.s.compare(0, 11, "DeclareGrid") == 0) {
      // The declaration stays (physis_hip.h turns it into a typedef); a
      // point type's support code follows it.
      size_t close = MatchClose(t_, i + 1);
      size_t end = close + 1 < e && t_[close + 1].s == ";" ? close + 1 : close;
      for (size_t q = i; q <= end; ++q) add(t_[q].s);
      out += "\n";
      auto u = g_utypes.find("PSGrid" + std::string(1, tk.s[11]) + "D" + t_[i + 2].s);
      if (u != g_utypes.end()) out += EmitUtypeSupport(u->second);
      i = end;
      continue;
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && ParseGridNew(tk.s, &gt) && !gt.utype.empty()) {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      if ((int)args.size() != gt.rank) Die(tk.s + ": expected " + std::to_string(gt.rank) + " extents");
      std::string r = "__PSGridNewUtype_" + gt.utype + "(";
      for (int d = 0; d < 3; ++d)
        r += (d ? ", " : "") + (d < gt.rank ? "(" + RewriteHost(args[d].first, args[d].second) + ")" : std::string("1"));
      add(r + ")");
      i = close;
      continue;
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && ParseGridNew(tk.s, &gt)) {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      if ((int)args.size() < gt.rank) Die(tk.s + ": expected " + std::to_string(gt.rank) + " extents");
      std::string attr = (int)args.size() > gt.rank ? RewriteHost(args[gt.rank].first, args[gt.rank].second) : "0";
      std::string r = "__PSGridNewHip(sizeof(" + gt.elem + "), " + std::to_string(gt.rank) + ", " + attr;
      for (int d = 0; d < 3; ++d)
        r += ", " + (d < gt.rank ? "(" + RewriteHost(args[d].first, args[d].second) + ")" : std::string("1"));
      r += ")";
      add(r);
      i = close;
      continue;
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && (tk.s == "PSReduce" || tk.s == "PSGridReduce")) {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      if (args.size() >= 4) {
        // PSReduce(&v, op, kernel, dom, args...): a reduction kernel.
        const std::string kn = Text(t_, args[2].first, args[2].second);
        if (!rkernels_.count(kn)) Die("PSReduce: '" + kn + "' is not a reduction kernel of this file");
        const Kernel &rk = rkernels_.at(kn);
        if (args.size() != 4 + rk.all.size())
          Die("PSReduce: kernel " + kn + " takes " + std::to_string(rk.all.size()) + " arguments after the domain");
        std::string call = "__PSReduceMap_" + kn + "(";
        for (size_t q = 0; q < args.size(); ++q) {
          if (q == 2) continue;
          call += (q ? ", " : "") + RewriteHost(args[q].first, args[q].second);
        }
        add(call + ")");
        i = close;
        continue;
      }
      if (args.size() != 3) Die("PSReduce takes a result address, an operator and a grid (or a kernel and a domain)");
      std::string g = Text(t_, args[2].first, args[2].second);
      auto it = gridvars_.find(g);
      if (it == gridvars_.end()) Die("PSReduce: unknown grid '" + g + "'");
      std::string fn;
      if (it->second.elem == "float") fn = "__PSReduceGridFloat";
      else if (it->second.elem == "double") fn = "__PSReduceGridDouble";
      else Die("PSReduce: unsupported element type of '" + g + "'");
      add(fn + "(" + RewriteHost(args[0].first, args[0].second) + ", " +
          RewriteHost(args[1].first, args[1].second) + ", " + g + ")");
      i = close;
      continue;
    }
    if (tk.kind == Tok::kIdent && i + 1 < e && t_[i + 1].s == "(" && tk.s == "PSStencilRun") {
      size_t close = MatchClose(t_, i + 1);
      auto args = SplitArgs(t_, i + 2, close);
      std::vector<std::pair<std::string, std::string>> maps;
      std::vector<int> modes;
      std::string iter = "1";
      static const std::map<std::string, int> kMapModes = {
          {"PSStencilMap", 0}, {"PSStencilMapRedBlack", 1}, {"PSStencilMapRed", 2}, {"PSStencilMapBlack", 3}};
      for (size_t a = 0; a < args.size(); ++a) {
        size_t ab = args[a].first, ae = args[a].second;
        auto mode = kMapModes.find(t_[ab].s);
        if (mode != kMapModes.end() && ab + 1 < ae && t_[ab + 1].s == "(") {
          modes.push_back(mode->second);
          size_t mc = MatchClose(t_, ab + 1);
          auto margs = SplitArgs(t_, ab + 2, mc);
          std::string kn = Text(t_, margs[0].first, margs[0].second);
          if (!kernels_.count(kn)) Die("PSStencilMap of unknown kernel '" + kn + "'");
          std::string call = "__PSStencilMap_" + kn + "(";
          for (size_t q = 1; q < margs.size(); ++q)
            call += (q > 1 ? ", " : "") + RewriteHost(margs[q].first, margs[q].second);
          call += ")";
          maps.push_back({kn, call});
        } else if (a + 1 == args.size()) {
          iter = RewriteHost(ab, ae);  // trailing count (reference run.cc:60-72)
        } else {
          Die("PSStencilRun arguments must be PSStencilMap(...) calls");
        }
      }
      int id = run_id_++;
      runs_.push_back(EmitRun(maps, modes, id));
      std::string call = "__PSStencilRun_" + std::to_string(id) + "(" + iter;
      for (auto &m : maps) call += ", " + m.second;
      call += ")";
      add(call);
      i = close;
      continue;
    }
    if (tk.kind == Tok::kPunct && tk.s == "(") ++pdepth;
    if (tk.kind == Tok::kPunct && tk.s == ")") --pdepth;
    if (tk.kind == Tok::kPunct && ((tk.s == ";" && pdepth == 0) || tk.s == "{" || tk.s == "}")) {
      add(tk.s);
      out += "\n";
      continue;
    }
    add(tk.s);
  }
  return out;
}

std::string Translator::Run(const std::vector<std::string> &includes) {
  CollectUtypes(t_);
  funcs_ = FindFunctions(t_);
  CollectGridVars();
  std::map<std::string, const Func *> byname;
  for (const Func &f : funcs_) byname[f.name] = &f;
  // Kernels: first argument of every PSStencilMap.
  for (size_t i = 0; i + 2 < t_.size(); ++i) {
    if (t_[i].kind == Tok::kIdent && t_[i + 1].s == "(" &&
        (t_[i].s == "PSStencilMap" || t_[i].s == "PSStencilMapRedBlack" || t_[i].s == "PSStencilMapRed" ||
         t_[i].s == "PSStencilMapBlack")) {
      const std::string &kn = t_[i + 2].s;
      auto it = byname.find(kn);
      if (it == byname.end()) Die("kernel '" + kn + "' is not defined in this file");
      if (!kernels_.count(kn)) {
        Kernel k;
        k.f = it->second;
        AnalyzeKernel(k);
        kernels_[kn] = k;
      }
    }
  }
  // Reduction kernels: third argument of every PSReduce with a domain.
  for (size_t i = 0; i + 2 < t_.size(); ++i) {
    if (t_[i].kind != Tok::kIdent || (t_[i].s != "PSReduce" && t_[i].s != "PSGridReduce") || t_[i + 1].s != "(")
      continue;
    const size_t close = MatchClose(t_, i + 1);
    const auto args = SplitArgs(t_, i + 2, close);
    if (args.size() < 4) continue;
    const std::string kn = Text(t_, args[2].first, args[2].second);
    if (rkernels_.count(kn)) continue;
    auto it = byname.find(kn);
    if (it == byname.end()) Die("reduction kernel '" + kn + "' is not defined in this file");
    if (kernels_.count(kn)) Die("kernel " + kn + " is used both as a stencil and as a reduction kernel");
    const Func &f = *it->second;
    const std::string ret = f.lparen >= 2 ? t_[f.lparen - 2].s : "";
    if (ret != "float" && ret != "double")
      Die("reduction kernel " + kn + " must return float or double (returns '" + ret + "')");
    Kernel k;
    k.f = it->second;
    AnalyzeKernel(k);
    if (k.emit_count) Die("reduction kernel " + kn + " must return its value, not emit it");
    // Every grid is read straight from HBM (no staged tile).
    k.staged.clear();
    k.rx = k.ry = k.rz = 0;
    k.wrap = false;
    rkernels_[kn] = k;
  }
  // Emit: header, user code with kernels replaced by the generated code
  // (inserted where the first kernel was defined).
  std::ostringstream o;
  o << "// Generated by physisc-hip from " << input_ << ". Do not edit.\n";
  o << "#define PHYSIS_HIP 1\n";
  for (const std::string &inc : includes) o << inc << "\n";
  o << "#include \"physis/physis.h\"\n#include \"physis/hip/generated.h\"\n\n";
  size_t first_kernel = t_.size();
  std::set<size_t> kernel_starts;
  for (auto *ks : {&kernels_, &rkernels_})
    for (auto &kv : *ks) {
      first_kernel = std::min(first_kernel, kv.second.f->begin);
      kernel_starts.insert(kv.second.f->begin);
    }
  std::string generated;
  for (auto &kv : kernels_) generated += EmitPlan(kv.second) + EmitMapAndLaunch(kv.second);
  for (auto &kv : rkernels_) generated += EmitReduceKernel(kv.second);
  // Host code, skipping kernel definitions (and their prototypes are harmless).
  std::string host_before, host_after;
  size_t pos = 0;
  std::vector<std::pair<size_t, size_t>> skip;
  for (auto *ks : {&kernels_, &rkernels_})
    for (auto &kv : *ks) skip.push_back({kv.second.f->begin, kv.second.f->rbrace + 1});
  std::sort(skip.begin(), skip.end());
  std::string *dst = &host_before;
  for (auto &sk : skip) {
    *dst += RewriteHost(pos, sk.first);
    if (sk.first == first_kernel) dst = &host_after;
    pos = sk.second;
  }
  *dst += RewriteHost(pos, t_.size());
  // The generated templates and runs keep C++ linkage even where the
  // program's kernels sit inside an extern "C" block (a C file compiled as
  // C++ for a C++ harness, as the reference's benchmarks are).
  o << host_before << "\nextern \"C++\" {\n" << generated;
  for (const std::string &r : runs_) o << r;
  o << "}  // extern \"C++\"\n" << host_after << "\n";
  return o.str();
}

std::string ReadFile(const std::string &p) {
  std::ifstream in(p);
  if (!in) Die("cannot read " + p);
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

// Runs the system C preprocessor over `src` with every #include replaced by
// an empty stand-in, so only the program's own macros are expanded; returns
// the expanded text and the include lines (re-emitted verbatim).
std::string Preprocess(const std::string &path, const std::vector<std::string> &defs,
                       std::vector<std::string> *includes) {
  std::string src = ReadFile(path);
  std::string tmpl = "/tmp/physisc-hip-XXXXXX";
  std::vector<char> dir(tmpl.begin(), tmpl.end());
  dir.push_back('\0');
  if (!mkdtemp(dir.data())) Die("mkdtemp failed");
  std::string tmp(dir.data());
  std::istringstream lines(src);
  std::string line, stripped;
  while (std::getline(lines, line)) {
    size_t b = line.find_first_not_of(" \t");
    if (b != std::string::npos && line[b] == '#') {
      size_t d = line.find_first_not_of(" \t", b + 1);
      if (d != std::string::npos && line.compare(d, 7, "include") == 0) {
        size_t q1 = line.find_first_of("<\"", d + 7);
        if (q1 != std::string::npos) {
          char close = line[q1] == '<' ? '>' : '"';
          size_t q2 = line.find(close, q1 + 1);
          std::string name = line.substr(q1 + 1, q2 - q1 - 1);
          if (name != "physis/physis.h") includes->push_back(line.substr(b));
          // empty stand-in under tmp/
          std::string full = tmp + "/" + name;
          for (size_t s = full.find('/', tmp.size() + 1); s != std::string::npos; s = full.find('/', s + 1))
            mkdir(full.substr(0, s).c_str(), 0700);
          FILE *f = fopen(full.c_str(), "w");
          if (f) fclose(f);
        }
      }
    }
    stripped += line + "\n";
  }
  std::string in_path = tmp + "/input.c";
  {
    std::ofstream f(in_path);
    f << stripped;
  }
  // The generated file is compiled as HIP C++, so the program's own
  // `#ifdef __cplusplus` branches (extern "C" blocks of a module) are taken.
  std::string cmd = "gcc -E -P -nostdinc -DPHYSIS_USER -D__cplusplus=201703L -I" + tmp;
  for (const std::string &d : defs) cmd += " '-D" + d + "'";
  cmd += " " + in_path + " 2>&1";
  FILE *p = popen(cmd.c_str(), "r");
  if (!p) Die("cannot run the C preprocessor");
  std::string out;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, p)) > 0) out.append(buf, n);
  int rc = pclose(p);
  std::string rm = "rm -rf '" + tmp + "'";
  if (system(rm.c_str()) != 0) fprintf(stderr, "physisc-hip: warning: could not remove %s\n", tmp.c_str());
  if (rc != 0) Die("preprocessing failed:\n" + out);
  return out;
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, output, config;
  std::vector<std::string> defs;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-o" && i + 1 < argc) output = argv[++i];
    else if (a == "--config" && i + 1 < argc) config = argv[++i];
    else if (a.compare(0, 2, "-D") == 0) defs.push_back(a.size() > 2 ? a.substr(2) : argv[++i]);
    else if (a == "--hip") continue;
    else if (a == "-h" || a == "--help") {
      printf("usage: physisc-hip [--config FILE] [-o OUT] [-D NAME[=V]]... INPUT.c\n");
      return 0;
    } else if (!a.empty() && a[0] == '-') Die("unknown option " + a);
    else input = a;
  }
  if (input.empty()) Die("no input file");
  if (output.empty()) {
    std::string base = input.substr(input.find_last_of('/') + 1);
    size_t dot = base.rfind('.');
    output = (dot == std::string::npos ? base : base.substr(0, dot)) + ".hip";
  }
  std::vector<std::string> includes;
  std::string expanded = Preprocess(input, defs, &includes);
  Translator tr(Tokenize(expanded), LoadConfig(config), input);
  std::string code = tr.Run(includes);
  std::ofstream out(output);
  if (!out) Die("cannot write " + output);
  out << code;
  return 0;
}
