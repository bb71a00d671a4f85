// Seeded synthetic workloads for the matching path.
//
//   csm_tool gen <c1|c2|c3|c4|c5|tiny> <out_dir> [--seed S] [--scale F]
//
// writes <out_dir>/data.graph, <out_dir>/update.stream, <out_dir>/q<k>.graph and
// <out_dir>/meta.txt ("key value" lines: n, m_initial, stream, inserts, deletes, batch_size, queries).
// Graph files use the reference's v/e format (README "Input File Format"); deletions in the update
// stream are "-e a b label" lines.  Stream construction guarantees the preconditions of the
// deletion extension: every deleted edge exists when its batch starts, inserted edges are new, and
// no edge is both deleted and inserted.
#pragma once

namespace csm {
int gen_main(int argc, char** argv);
}
