"""CPU FP64 oracle of QSF's split-operator path. Test infrastructure only."""
