// regex.cpp -- compiles the regular expressions the reference hands to
// libcudf's contains_re (the LIKE translation of LogicalProject.cpp:31-47,
// searched at :113-131) into a state machine the device matcher runs by
// bitsets (interp_kernels.hip, regex_contains).
//
// Syntax: literals (UTF-8), '.', classes [..] / [^..] with ranges and \d \w
// \s inside, escapes \d \w \s \D \W \S \n \t \r \f \v and escaped
// punctuation, groups (..) and (?:..), alternation |, quantifiers * + ?
// {n} {n,} {n,m} (and their lazy forms, the same for a search), anchors ^
// (start of the string) and $ (end of the string). '.' matches any character
// but '\n'. Anything else (backreferences, lookaround, \b, a stray '{') is
// rejected with a message.
//
// Layout of the compiled program (int32 words), P states, state P-1 = MATCH:
//   [0] P   [1] offset of the class data   [2] offset of the edge table
//   [3] capture groups (group 0 = the whole match not counted)
//   [4 + q]            state q: kind (bits 0-2: 0 char, 1 any, 2 class,
//                      3 match, 4 epsilon) | argument << 3 (code point or
//                      class offset)
//   [4 + P + 4q]       closure of state q in the middle of the string (128-bit
//                      set of char/any/class/match states, 4 words)
//   [4 + 5P + 4q]      closure of state q at the end of the string ($ passes)
//   [4 + 9P + 4k]      closure of state 0 with k = 2*(at start) + (at end)
//   classes: n_ranges, negate, then n_ranges (lo, hi) code point pairs
//   edges, 2 words per state (for the backtracking matcher): x | y << 16
//   (0xFFFF: none; x is tried first) and cond | (save slot + 1) << 8 (cond
//   1: only at the start, 2: only at the end; slot 2g / 2g+1 = start / end
//   of group g, g >= 1; cond bit 7: the split of an unbounded loop, whose
//   body starts at the next state and whose last body state jumps back to
//   it).
#include <cstdint>
#include <string>
#include <vector>

#include "program.hpp"

namespace bsql {
namespace {

constexpr int kMaxStates = 128;

enum Kind { K_CHAR = 0, K_ANY = 1, K_CLASS = 2, K_MATCH = 3, K_EPS = 4 };

struct Node {
  enum T { EMPTY, CHAR, ANY, CLASS, BOL, EOL, CAT, ALT, REP } t = EMPTY;
  uint32_t cp = 0;
  std::vector<std::pair<uint32_t, uint32_t>> ranges;
  bool negate = false;
  std::vector<Node> kids;
  int min = 0, max = 0;  // REP; max < 0: unbounded
  bool lazy = false;     // REP
  int group = 0;         // CAT: capture group index (0: none)
};

struct Parser {
  const std::string& s;
  size_t i = 0;
  int groups = 0;
  explicit Parser(const std::string& re) : s(re) {}

  [[noreturn]] void bad(const std::string& why) { fail("regex '" + s + "': " + why); }

  uint32_t utf8() {
    const unsigned char c = (unsigned char)s[i];
    int n = c < 0x80 ? 1 : c < 0xE0 ? 2 : c < 0xF0 ? 3 : 4;
    if (i + (size_t)n > s.size()) bad("invalid UTF-8");
    uint32_t cp = n == 1 ? c : n == 2 ? (c & 0x1F) : n == 3 ? (c & 0x0F) : (c & 0x07);
    for (int k = 1; k < n; ++k) cp = (cp << 6) | ((unsigned char)s[i + (size_t)k] & 0x3F);
    i += (size_t)n;
    return cp;
  }

  static void shorthand(char c, Node& n) {
    auto add = [&](uint32_t lo, uint32_t hi) { n.ranges.push_back({lo, hi}); };
    switch (c) {
      case 'd': case 'D': add('0', '9'); break;
      case 'w': case 'W': add('0', '9'); add('A', 'Z'); add('_', '_'); add('a', 'z'); break;
      default: add('\t', '\r'); add(' ', ' '); break;  // s, S: \t \n \v \f \r and space
    }
  }

  // an escape after '\': a class shorthand (into n) or one code point
  bool escape(Node& n, uint32_t* cp) {
    if (i >= s.size()) bad("trailing backslash");
    const char c = s[i++];
    switch (c) {
      case 'd': case 'w': case 's': shorthand(c, n); return true;
      case 'D': case 'W': case 'S': shorthand(c, n); n.negate = !n.negate; return true;
      case 'n': *cp = '\n'; return false;
      case 't': *cp = '\t'; return false;
      case 'r': *cp = '\r'; return false;
      case 'f': *cp = '\f'; return false;
      case 'v': *cp = '\v'; return false;
      default:
        if ((c >= '0' && c <= '9') || (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'))
          bad(std::string("escape \\") + c + " is not supported");
        *cp = (unsigned char)c;
        return false;
    }
  }

  Node cls() {  // after '['
    Node n;
    n.t = Node::CLASS;
    if (i < s.size() && s[i] == '^') { n.negate = true; ++i; }
    if (i < s.size() && s[i] == ']') bad("empty class");
    while (true) {
      if (i >= s.size()) bad("unterminated class");
      if (s[i] == ']') { ++i; break; }
      uint32_t lo;
      if (s[i] == '\\') {
        ++i;
        Node sh;
        if (escape(sh, &lo)) {
          if (sh.negate) bad("negated shorthand inside a class is not supported");
          n.ranges.insert(n.ranges.end(), sh.ranges.begin(), sh.ranges.end());
          continue;
        }
      } else {
        lo = utf8();
      }
      uint32_t hi = lo;
      if (i + 1 < s.size() && s[i] == '-' && s[i + 1] != ']') {
        ++i;
        if (s[i] == '\\') {
          ++i;
          Node sh;
          if (escape(sh, &hi)) bad("class range ending in a shorthand");
        } else {
          hi = utf8();
        }
        if (hi < lo) bad("class range out of order");
      }
      n.ranges.push_back({lo, hi});
    }
    return n;
  }

  Node atom() {
    Node n;
    const char c = s[i];
    if (c == '(') {
      ++i;
      int g = 0;
      if (s.compare(i, 2, "?:") == 0) i += 2;
      else if (i < s.size() && s[i] == '?') bad("lookaround is not supported");
      else if (++groups > 9) bad("more than 9 capture groups");
      else g = groups;
      Node inner = alt();
      if (i >= s.size() || s[i] != ')') bad("unbalanced parenthesis");
      ++i;
      if (!g) return inner;
      n.t = Node::CAT;
      n.group = g;
      n.kids.push_back(inner);
      return n;
    }
    if (c == '[') { ++i; return cls(); }
    if (c == '.') { ++i; n.t = Node::ANY; return n; }
    if (c == '^') { ++i; n.t = Node::BOL; return n; }
    if (c == '$') { ++i; n.t = Node::EOL; return n; }
    if (c == '\\') {
      ++i;
      uint32_t cp;
      Node sh;
      if (escape(sh, &cp)) { sh.t = Node::CLASS; return sh; }
      n.t = Node::CHAR;
      n.cp = cp;
      return n;
    }
    if (c == '*' || c == '+' || c == '?') bad("quantifier without an operand");
    if (c == '{' || c == '}' || c == ']') bad(std::string("unexpected '") + c + "'");
    n.t = Node::CHAR;
    n.cp = utf8();
    return n;
  }

  int number() {
    if (i >= s.size() || s[i] < '0' || s[i] > '9') bad("malformed {m,n}");
    int v = 0;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
      v = v * 10 + (s[i++] - '0');
      if (v > 1000) bad("repeat count too large");
    }
    return v;
  }

  Node repeat() {
    Node n = atom();
    while (i < s.size() && (s[i] == '*' || s[i] == '+' || s[i] == '?' || s[i] == '{')) {
      Node r;
      r.t = Node::REP;
      const char q = s[i++];
      if (q == '*') { r.min = 0; r.max = -1; }
      else if (q == '+') { r.min = 1; r.max = -1; }
      else if (q == '?') { r.min = 0; r.max = 1; }
      else {
        r.min = number();
        r.max = r.min;
        if (i < s.size() && s[i] == ',') {
          ++i;
          r.max = (i < s.size() && s[i] == '}') ? -1 : number();
        }
        if (i >= s.size() || s[i] != '}') bad("malformed {m,n}");
        ++i;
        if (r.max >= 0 && r.max < r.min) bad("{m,n} with n < m");
      }
      if (i < s.size() && s[i] == '?') {  // lazy: the same matches for a search, fewer for a replace
        r.lazy = true;
        ++i;
      }
      if (n.t == Node::BOL || n.t == Node::EOL) bad("quantified anchor");
      r.kids.push_back(n);
      n = r;
    }
    return n;
  }

  Node cat() {
    Node n;
    n.t = Node::CAT;
    while (i < s.size() && s[i] != '|' && s[i] != ')') n.kids.push_back(repeat());
    return n;
  }

  Node alt() {
    Node first = cat();
    if (i >= s.size() || s[i] != '|') return first;
    Node n;
    n.t = Node::ALT;
    n.kids.push_back(first);
    while (i < s.size() && s[i] == '|') {
      ++i;
      n.kids.push_back(cat());
    }
    return n;
  }
};

struct State {
  int kind;
  uint32_t arg = 0;                  // code point or index into classes
  int x = -1, y = -1;                // epsilon edges (-1: none; consuming states go to +1); x first
  int cond = 0;                      // epsilon condition: 1 start only, 2 end only; 128: loop split
  int save = -1;                     // capture slot written when passing
};

struct Builder {
  std::vector<State> st;
  std::vector<const Node*> classes;

  int add(State s) {
    st.push_back(s);
    if ((int)st.size() > kMaxStates) fail("regex too large for the device matcher");
    return (int)st.size() - 1;
  }
  void emit(const Node& n) {
    switch (n.t) {
      case Node::EMPTY: return;
      case Node::CHAR: add({K_CHAR, n.cp}); return;
      case Node::ANY: add({K_ANY}); return;
      case Node::CLASS: {
        classes.push_back(&n);
        add({K_CLASS, (uint32_t)classes.size() - 1});
        return;
      }
      case Node::BOL: case Node::EOL: {
        State e{K_EPS};
        e.cond = n.t == Node::BOL ? 1 : 2;
        const int q = add(e);
        st[(size_t)q].x = q + 1;
        return;
      }
      case Node::CAT: {
        if (n.group) {
          State e{K_EPS};
          e.save = 2 * n.group;
          const int q = add(e);
          st[(size_t)q].x = q + 1;
        }
        for (const Node& k : n.kids) emit(k);
        if (n.group) {
          State e{K_EPS};
          e.save = 2 * n.group + 1;
          const int q = add(e);
          st[(size_t)q].x = q + 1;
        }
        return;
      }
      case Node::ALT: {
        std::vector<int> jumps;
        for (size_t k = 0; k < n.kids.size(); ++k) {
          if (k + 1 < n.kids.size()) {
            const int split = add({K_EPS});
            st[(size_t)split].x = split + 1;
            emit(n.kids[k]);
            jumps.push_back(add({K_EPS}));
            st[(size_t)split].y = (int)st.size();
          } else {
            emit(n.kids[k]);
          }
        }
        for (int j : jumps) st[(size_t)j].x = (int)st.size();
        return;
      }
      case Node::REP: {
        const Node& body = n.kids[0];
        for (int k = 0; k < n.min; ++k) emit(body);
        // a split tries the body first (greedy) or the way out first (lazy)
        auto order = [&](int split, int body_at, int out_at) {
          st[(size_t)split].x = n.lazy ? out_at : body_at;
          st[(size_t)split].y = n.lazy ? body_at : out_at;
        };
        if (n.max < 0) {
          const int split = add({K_EPS});
          st[(size_t)split].cond = 128;  // the matcher repeats no iteration that consumed nothing
          emit(body);
          const int back = add({K_EPS});
          st[(size_t)back].x = split;
          order(split, split + 1, (int)st.size());
        } else {
          std::vector<int> splits;
          for (int k = n.min; k < n.max; ++k) {
            const int split = add({K_EPS});
            splits.push_back(split);
            emit(body);
          }
          for (int sp : splits) order(sp, sp + 1, (int)st.size());
        }
        return;
      }
    }
  }
};

// states reachable from q by epsilon edges, as a 128-bit set of the
// non-epsilon states among them
void closure(const std::vector<State>& st, int q, bool at_start, bool at_end, uint32_t out[4]) {
  std::vector<char> seen(st.size(), 0);
  std::vector<int> stack{q};
  out[0] = out[1] = out[2] = out[3] = 0;
  while (!stack.empty()) {
    const int p = stack.back();
    stack.pop_back();
    if (p < 0 || p >= (int)st.size() || seen[(size_t)p]) continue;
    seen[(size_t)p] = 1;
    const State& s = st[(size_t)p];
    if (s.kind != K_EPS) {
      out[p >> 5] |= 1u << (p & 31);
      continue;
    }
    if ((s.cond == 1 && !at_start) || (s.cond == 2 && !at_end)) continue;
    stack.push_back(s.x);
    if (s.y >= 0) stack.push_back(s.y);
  }
}

}  // namespace

std::vector<int32_t> regex_compile(const std::string& re) {
  Parser p(re);
  Node root = p.alt();
  if (p.i != re.size()) p.bad("unbalanced parenthesis");
  Builder b;
  b.emit(root);
  b.add({K_MATCH});
  const int P = (int)b.st.size();
  std::vector<int32_t> out((size_t)(4 + P + 8 * P + 16), 0);
  out[0] = P;
  out[3] = p.groups;
  std::vector<int32_t> cls;
  std::vector<int32_t> class_at;
  for (const Node* c : b.classes) {
    class_at.push_back((int32_t)cls.size());
    cls.push_back((int32_t)c->ranges.size());
    cls.push_back(c->negate ? 1 : 0);
    for (const auto& r : c->ranges) {
      cls.push_back((int32_t)r.first);
      cls.push_back((int32_t)r.second);
    }
  }
  const int32_t class_base = (int32_t)out.size();
  out[1] = class_base;
  for (int q = 0; q < P; ++q) {
    const State& s = b.st[(size_t)q];
    uint32_t arg = s.kind == K_CLASS ? (uint32_t)(class_base + class_at[s.arg]) : s.arg;
    out[(size_t)(4 + q)] = (int32_t)((uint32_t)s.kind | (arg << 3));
    uint32_t m[4];
    closure(b.st, q, false, false, m);
    for (int w = 0; w < 4; ++w) out[(size_t)(4 + P + 4 * q + w)] = (int32_t)m[w];
    closure(b.st, q, false, true, m);
    for (int w = 0; w < 4; ++w) out[(size_t)(4 + 5 * P + 4 * q + w)] = (int32_t)m[w];
  }
  for (int k = 0; k < 4; ++k) {
    uint32_t m[4];
    closure(b.st, 0, (k & 2) != 0, (k & 1) != 0, m);
    for (int w = 0; w < 4; ++w) out[(size_t)(4 + 9 * P + 4 * k + w)] = (int32_t)m[w];
  }
  out.insert(out.end(), cls.begin(), cls.end());
  out[2] = (int32_t)out.size();
  for (int q = 0; q < P; ++q) {
    const State& s = b.st[(size_t)q];
    const uint32_t x = s.x < 0 ? 0xFFFFu : (uint32_t)s.x, y = s.y < 0 ? 0xFFFFu : (uint32_t)s.y;
    out.push_back((int32_t)(x | (y << 16)));
    out.push_back((int32_t)((uint32_t)s.cond | ((uint32_t)(s.save + 1) << 8)));
  }
  return out;
}

}  // namespace bsql
