"""MI355X-native engine for MarginPolish's realignment hot path: the banded
pair-HMM forward-backward with run-length-aware emissions and posterior
match/insert/delete extraction (reference: impl/pairwiseAligner.c,
impl/stateMachine.c, impl/polisher.c:605-703)."""
__version__ = "0.1.0"
