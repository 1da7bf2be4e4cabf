"""The oracle: a CPU restatement of the reference hot path, used only as the
checker by tests/, the smoke check and the bench's CPU baseline."""
