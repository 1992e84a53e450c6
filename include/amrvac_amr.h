/* amrvac_amr.h -- native host routines of the block-structured AMR tree.
 *
 * The tree of a level hierarchy is given by its leaves: level (1-based) and
 * integer block position ig (1-based, per dimension, on that level's grid of
 * ng1 * 2^(level-1) blocks), one entry per leaf in block order.  A node
 * (level, ig) exists when some leaf lies at or below it; it is a leaf when it
 * is one of the given entries.  Directions i in {-1, 0, 1}^ndim are indexed
 * k = sum (i_d + 1) 3^d (dimension 1 fastest), children j = sum (ic_d - 1) 2^d.
 * Neighbour types (find_neighbor src/amr/mod_amr_neighbors.t:54-124):
 * 0 the block itself, 1 physical boundary, 2 coarser, 3 same level, 4 finer.
 *
 * Both routines rebuild a hash of the nodes from the leaves (no state is kept
 * between calls) and return 0, or a nonzero code with a message from
 * amrvac_amr_last_error().
 */
#ifndef AMRVAC_AMR_H
#define AMRVAC_AMR_H

#ifdef __cplusplus
extern "C" {
#endif

/* build_connectivity src/mod_functions_connectivity.t:55-431 (one rank):
 * per leaf b and direction k, ntype[b*3^nd + k], nbr[b*3^nd + k] (the
 * same-level or coarser leaf, the block itself for k's centre, else -1) and
 * nchild[(b*3^nd + k)*2^nd + j] (for a finer neighbour: its children j that
 * touch block b, else -1).  periodic[d] wraps direction d.  Returns 2 when a
 * neighbour is more than one level finer. */
int amrvac_amr_connectivity(int ndim, const int* ng1, const int* periodic, int nleaves, const int* level,
                            const int* ig, int* ntype, int* nbr, int* nchild);

/* proper_nesting src/amr/mod_coarsen_refine.t:186-319: 2:1 balance of the
 * refine / coarsen flags (one byte per leaf, updated in place).  Refining a
 * leaf refines the leaf neighbours of its parent that lie outward; a
 * coarsening needs every sibling flagged and no same-level neighbour
 * refining and no finer neighbour staying finer, else the whole family is
 * unflagged.  order[] lists the leaves in the order the coarsening checks
 * visit them within a level (the level lists of the tree). */
int amrvac_amr_proper_nesting(int ndim, const int* ng1, const int* periodic, int nleaves, const int* level,
                              const int* ig, const int* order, int refine_max_level, unsigned char* refine,
                              unsigned char* coarsen);

const char* amrvac_amr_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AMRVAC_AMR_H */
